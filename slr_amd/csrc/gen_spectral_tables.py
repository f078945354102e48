"""Regenerates spectral_tables.inc, the two published data tables the
spectral path needs bit-for-bit, and spectral_library.inc, the illuminant
spectra scene files can name:

* the CIE 1931 2-degree standard observer, 360-830 nm at 1 nm (471 samples),
  as tabulated in libSLR/BasicTypes/spectrum_base.cpp:35-221;
* the Meng et al. 2015 ("Physically Meaningful Rendering using Tristimulus
  Colours") RGB-to-spectrum basis: a 12x14 grid of cells over the (u, v)
  chromaticity plane and its data points, each a 95-sample spectrum at 5 nm
  from 360 to 830 nm (libSLR/BasicTypes/spectrum_types.cpp:1004-1365).

Both are data, not code: the renderer's colours depend on every value, so
the tables are carried over exactly (rounded to float32 the way the
reference's float instantiation stores them) into this repository's own
layout. Run once, with the reference checkout at REF:

    python slr_amd/csrc/gen_spectral_tables.py [REF]
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _block(text, start_pat, end_pat="};"):
    i = text.index(start_pat)
    j = text.index(end_pat, i)
    return text[i + len(start_pat):j]


def _floats(s):
    return [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", s)]


def main():
    base = open(os.path.join(REF, "libSLR/BasicTypes/spectrum_base.cpp")).read()
    cmf = []
    for name in ("xbar_CIE1931_2deg[] = {", "ybar_CIE1931_2deg[] = {", "zbar_CIE1931_2deg[] = {"):
        v = _floats(_block(base, name))
        assert len(v) == 471, (name, len(v))
        cmf.append(np.float32(v))

    st = open(os.path.join(REF, "libSLR/BasicTypes/spectrum_types.cpp")).read()
    grid_txt = _block(st, "UpsampledContinuousSpectrumTemplate<float, NumSpectralSamples>::spectrum_grid[] = {")
    cells = []
    for m in re.finditer(r"\{\s*(\d+),\s*(\d+),\s*\{([^}]*)\}\s*\}", grid_txt):
        idx = [255 if t.strip() == "UINT8_MAX" else int(t) for t in m.group(3).split(",")]
        cells.append((int(m.group(1)), int(m.group(2)), idx))
    assert len(cells) == 12 * 14, len(cells)

    pts_txt = _block(st, "UpsampledContinuousSpectrumTemplate<float, NumSpectralSamples>::spectrum_data_points[] = {",
                     "\n    };")
    points = []
    for m in re.finditer(r"\{\s*\{([^}]*)\},\s*\{([^}]*)\},\s*\{([^}]*)\}\s*\}", pts_txt):
        uv = _floats(m.group(2))
        sp = _floats(m.group(3))
        assert len(uv) == 2 and len(sp) == 95
        points.append((np.float32(uv), np.float32(sp)))
    out = []
    w = out.append
    w("// Generated by gen_spectral_tables.py; see that script for the provenance")
    w("// of these published tables. Do not edit by hand.")
    w("#pragma once")
    w("#include <cstdint>")
    w("namespace slrh { namespace tables {")
    w("constexpr int kNumCMFSamples = 471;  // 360..830 nm, 1 nm")
    for nm, arr in zip("XYZ", cmf):
        w("alignas(16) const float kCIE1931_%s[kNumCMFSamples] = {" % nm)
        for i in range(0, 471, 8):
            w("  " + ", ".join("%.9g" % x for x in arr[i:i + 8]) + ",")
        w("};")
    w("constexpr int kMengGridW = 12, kMengGridH = 14;")
    w("constexpr int kMengNumPoints = %d;" % len(points))
    w("constexpr int kMengNumLambdas = 95;  // 360..830 nm, 5 nm")
    w("struct MengCell { uint8_t inside, numPoints, idx[6]; };")
    w("const MengCell kMengCells[kMengGridW * kMengGridH] = {")
    for ins, npt, idx in cells:
        w("  {%d, %d, {%s}}," % (ins, npt, ", ".join(str(i) for i in idx)))
    w("};")
    w("alignas(16) const float kMengUV[kMengNumPoints][2] = {")
    for uv, _ in points:
        w("  {%.9g, %.9g}," % (uv[0], uv[1]))
    w("};")
    w("alignas(16) const float kMengSpectra[kMengNumPoints][kMengNumLambdas] = {")
    for _, sp in points:
        w("  {" + ", ".join("%.9g" % x for x in sp) + "},")
    w("};")
    w("}}  // namespace slrh::tables")
    with open(os.path.join(HERE, "spectral_tables.inc"), "w") as f:
        f.write("\n".join(out) + "\n")
    print("points", len(points), "cells", len(cells))
    illuminants()


def illuminants():
    """spectral_library.inc: the illuminant spectra a scene file can name
    (Spectrum("Illuminant", name, 0)), from the regular tables of
    libSLR/BasicTypes/spectrum_library.cpp:12-103 (CIE D65 300-830 nm at
    1 nm, and E scaled to D65's energy)."""
    lib = open(os.path.join(REF, "libSLR/BasicTypes/spectrum_library.cpp"), encoding="utf-8-sig").read()
    out = []
    w = out.append
    w("// Generated by gen_spectral_tables.py from the reference's spectrum library")
    w("// (libSLR/BasicTypes/spectrum_library.cpp:12-103): data, not code.")
    w("#pragma once")
    w("#include <cstdint>")
    w("namespace slrh { namespace tables {")
    w("struct LibrarySpectrum { const char* name; uint32_t numSamples; float minLambda, maxLambda; const float* values; };")
    entries = []
    for name in ("D65", "E"):
        vals = np.float32(_floats(_block(lib, "const float %s_values[] = {" % name)))
        m = re.search(r"const Data %s = \{\s*DistributionType::Regular,\s*(\d+),\s*([\d.]+)f,\s*([\d.]+)f" % name, lib)
        n, lo, hi = int(m.group(1)), float(m.group(2)), float(m.group(3))
        assert n == len(vals), (name, n, len(vals))
        w("alignas(16) const float kIlluminant%s[%d] = {" % (name, n))
        for i in range(0, n, 8):
            w("  " + ", ".join("%.9g" % x for x in vals[i:i + 8]) + ",")
        w("};")
        entries.append('{"%s", %d, %.1ff, %.1ff, kIlluminant%s}' % (name, n, lo, hi, name))
    w("const LibrarySpectrum kIlluminants[] = {" + ", ".join(entries) + "};")
    w("}}  // namespace slrh::tables")
    with open(os.path.join(HERE, "spectral_library.inc"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()

// Reader for a subset of SLR's scene language.
//
// The language (libSLRSceneGraph/Parser/SceneLexer.l, SceneParser.yy) has
// statements `expr;`, variables, integer/real/bool/string literals, tuples
// with unnamed and named ("key": value) elements, arithmetic, indexing and
// calls of builtin functions. A call maps its arguments to one of the
// function's signatures the way mapParamsToArgs does
// (libSLRSceneGraph/Parser/SceneParser.cpp:387-463): named arguments by key,
// then each remaining parameter takes the next unnamed argument if that is
// convertible to its type, else its default; the first signature that maps
// wins. Supported here: the builtins a scene such as
// TestScenes/Original_Cornell_Box.txt uses (API.cpp:295-1325) — Spectrum
// (five forms), SpectrumTexture, FloatTexture, createSurfaceMaterial
// ("matte", "microfacet metal", "microfacet glass", "emitter"),
// createEmitterSurfaceProperty ("diffuse"), createNode, createMesh,
// createVertex, addChild, setTransform, translate, scale, rotate,
// rotateX/Y/Z, createPerspectiveCamera, setRenderer ("PT",
// "Volumetric PT"), setRenderSettings, and the predefined `root`. Control
// flow (if, for, function), image and model loading and the other materials
// are reported as outside the subset.
//
// Scene assembly follows libSLR: a mesh's vertex normals are the normalised
// sums of its face normals and its tangents makeCoordinateSystem's
// (TriangleMeshNode::recalculateNormalAndTangents,
// libSLRSceneGraph/Scene/TriangleMeshNode.cpp:106-134); a node's transform
// is its parent's times its own (ChainedTransform::reduce,
// libSLR/Core/transform.cpp:25-84), in FP32 as Matrix4x4 multiplies; mesh
// vertices go to world space with it, normals with its inverse transposed
// and renormalised (libSLR/Scene/TriangleMeshNode.cpp:61-75); the camera
// takes its node's transform.
#include "scene_reader.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>

#include "spectral_library.inc"

namespace slrh {
namespace scenefile {

namespace {

// ------------------------------------------------------------------ values

enum class Ty {
    Void, Bool, Integer, Real, String, Tuple, Spectrum, SpectrumTexture, FloatTexture, SurfaceMaterial,
    EmitterProperty, InternalNode, SurfaceNode, CameraNode, Matrix, Vertex,
    NormalTexture,   // parameter type only: nothing here makes one
    Node,        // parameter type: any node
    Transform,   // parameter type: a matrix
    Any
};

struct Mat4 {
    float m[4][4];   // [column][row], as Matrix4x4's c0..c3
    static Mat4 identity() {
        Mat4 r;
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) r.m[c][k] = c == k ? 1.0f : 0.0f;
        return r;
    }
    float at(int row, int col) const { return m[col][row]; }
};

// Matrix4x4::operator*(Matrix4x4) (libSLR/BasicTypes/Matrix4x4.h:57-63):
// each entry a row of the left dotted with a column of the right, in order.
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r.m[c][row] = a.at(row, 0) * b.m[c][0] + a.at(row, 1) * b.m[c][1] + a.at(row, 2) * b.m[c][2] +
                          a.at(row, 3) * b.m[c][3];
    return r;
}

struct Obj;
struct Value;
struct TupleV {
    std::vector<Value> unnamed;
    std::vector<std::pair<std::string, Value>> named;
};

struct Value {
    Ty ty = Ty::Void;
    bool b = false;
    int32_t i = 0;
    double r = 0;
    std::string s;
    std::shared_ptr<TupleV> tup;
    std::shared_ptr<Obj> obj;
    Mat4 mat = Mat4::identity();
    double real() const { return ty == Ty::Real ? r : ty == Ty::Integer ? (double)i : (double)b; }
};

struct Node;
struct Obj {
    int index = -1;                 // spectrum / material index in the SceneFile
    std::shared_ptr<Node> node;     // node objects
    int spectrum = -1;              // texture: its spectrum
    float value = 0;                // float texture
    float vtx[3] = {0, 0, 0};       // vertex position
};

struct Node {
    enum Kind { kInternalNode, kMeshNode, kCameraNode } kind = kInternalNode;
    Mat4 local = Mat4::identity();
    std::vector<std::shared_ptr<Node>> children;
    // mesh
    std::vector<float> pos;                      // object space, 3 per vertex
    std::vector<uint32_t> idx, mat;              // 3 per triangle, material per triangle
    // camera
    Camera cam;
};

Value mkInt(int32_t v) { Value x; x.ty = Ty::Integer; x.i = v; return x; }
Value mkReal(double v) { Value x; x.ty = Ty::Real; x.r = v; return x; }
Value mkStr(const std::string& v) { Value x; x.ty = Ty::String; x.s = v; return x; }
Value mkTuple() { Value x; x.ty = Ty::Tuple; x.tup = std::make_shared<TupleV>(); return x; }
Value mkObj(Ty t, std::shared_ptr<Obj> o) { Value x; x.ty = t; x.obj = std::move(o); return x; }

const char* tyName(Ty t) {
    switch (t) {
        case Ty::Void: return "void";
        case Ty::Bool: return "bool";
        case Ty::Integer: return "integer";
        case Ty::Real: return "real number";
        case Ty::String: return "string";
        case Ty::Tuple: return "tuple";
        case Ty::Spectrum: return "spectrum";
        case Ty::SpectrumTexture: return "spectrum texture";
        case Ty::FloatTexture: return "float texture";
        case Ty::SurfaceMaterial: return "surface material";
        case Ty::EmitterProperty: return "emitter surface property";
        case Ty::InternalNode: return "internal node";
        case Ty::SurfaceNode: return "surface node";
        case Ty::CameraNode: return "camera node";
        case Ty::Matrix: return "matrix";
        case Ty::Vertex: return "vertex";
        default: return "value";
    }
}

// TypeInfo conversions (SceneParser.cpp:555-935): every type to itself and
// Any; bool and integer to each other and to real; nodes to Node; a matrix
// to Transform.
bool convertible(Ty from, Ty to) {
    if (from == to || to == Ty::Any) return from != Ty::Void;
    switch (from) {
        case Ty::Bool: return to == Ty::Integer || to == Ty::Real;
        case Ty::Integer: return to == Ty::Bool || to == Ty::Real;
        case Ty::InternalNode:
        case Ty::SurfaceNode:
        case Ty::CameraNode: return to == Ty::Node;
        case Ty::Matrix: return to == Ty::Transform;
        default: return false;
    }
}
Value convert(const Value& v, Ty to) {
    if (v.ty == to || to == Ty::Any || to == Ty::Node || to == Ty::Transform) return v;
    if (to == Ty::Real) return mkReal(v.real());
    if (to == Ty::Integer) return mkInt(v.ty == Ty::Bool ? (int32_t)v.b : v.i);
    if (to == Ty::Bool) { Value x; x.ty = Ty::Bool; x.b = v.ty == Ty::Integer ? v.i != 0 : v.b; return x; }
    return v;
}

// ------------------------------------------------------------------ builtin calls

struct Arg {
    std::string name;
    Ty type;
    bool hasDefault = false;
    Value def;
};
Arg req(const char* n, Ty t) { Arg a; a.name = n; a.type = t; return a; }
Arg opt(const char* n, Ty t, Value d) { Arg a; a.name = n; a.type = t; a.hasDefault = true; a.def = std::move(d); return a; }
Arg optNull(const char* n, Ty t) { Arg a; a.name = n; a.type = t; a.hasDefault = true; return a; }

typedef std::map<std::string, Value> Args;

// mapParamsToArgs (SceneParser.cpp:387-463).
bool mapArgs(const TupleV& params, const std::vector<Arg>& sig, Args* args) {
    args->clear();
    std::vector<bool> assigned(sig.size(), false);
    std::map<std::string, bool> namedUsed;
    for (auto& kv : params.named) namedUsed[kv.first] = false;
    for (size_t k = 0; k < sig.size(); ++k) {
        const Value* found = nullptr;
        for (auto& kv : params.named)
            if (kv.first == sig[k].name) found = &kv.second;
        if (!found) continue;
        if (!convertible(found->ty, sig[k].type)) return false;
        (*args)[sig[k].name] = convert(*found, sig[k].type);
        assigned[k] = true;
        namedUsed[sig[k].name] = true;
    }
    for (auto& kv : namedUsed)
        if (!kv.second) return false;
    size_t next = 0;
    for (size_t k = 0; k < sig.size(); ++k) {
        if (assigned[k]) continue;
        if (next < params.unnamed.size() && convertible(params.unnamed[next].ty, sig[k].type)) {
            (*args)[sig[k].name] = convert(params.unnamed[next], sig[k].type);
            ++next;
        } else {
            if (!sig[k].hasDefault) return false;
            (*args)[sig[k].name] = sig[k].def;
        }
        assigned[k] = true;
    }
    return next >= params.unnamed.size();
}

struct Interp;
typedef std::function<Value(Interp&, const Args&)> Proc;
struct Builtin {
    std::vector<std::vector<Arg>> sigs;
    std::vector<Proc> procs;
};

struct ScriptError {
    std::string msg;
};
[[noreturn]] void fail(const std::string& m) { throw ScriptError{m}; }

// ------------------------------------------------------------------ spectra

// IrregularContinuousSpectrum::evaluate at one wavelength
// (libSLR/BasicTypes/spectrum_types.cpp:286-303).
float irregularAt(const std::vector<float>& lambdas, const std::vector<float>& values, float wl) {
    const int32_t n = (int32_t)lambdas.size();
    const int32_t lowIdx =
        std::max((int32_t)(std::lower_bound(lambdas.begin(), lambdas.end(), wl) - lambdas.begin()) - 1, 0);
    if (lowIdx >= n - 1) return values[n - 1];
    const float t = (wl - lambdas[lowIdx]) / (lambdas[lowIdx + 1] - lambdas[lowIdx]);
    if (t <= 0.0f) return values[0];
    return (1 - t) * values[lowIdx] + t * values[lowIdx + 1];
}

// An irregularly sampled spectrum as a regular table over its first to last
// wavelength whose spacing is the largest that puts every sample wavelength
// on a knot (wavelengths in steps of 1/1000 nm): the same piecewise-linear
// function, its values at the knots from the reference's own evaluation.
Spectrum irregularAsRegular(int type, const std::vector<float>& lambdas, const std::vector<float>& values) {
    if (lambdas.size() < 2) {
        Spectrum s;
        s.kind = kRegular;
        s.type = type;
        s.lo = s.hi = lambdas.empty() ? 360.0f : lambdas[0];
        s.values = values.empty() ? std::vector<float>{0.0f} : std::vector<float>{values[0]};
        return s;
    }
    long long g = 0;
    std::vector<long long> milli(lambdas.size());
    for (size_t k = 0; k < lambdas.size(); ++k) {
        const double mm = std::round((double)lambdas[k] * 1000.0);
        if (std::fabs(mm - (double)lambdas[k] * 1000.0) > 1e-2) fail("spectrum wavelengths finer than 1/1000 nm");
        milli[k] = (long long)mm;
        if (k > 0) {
            if (milli[k] <= milli[k - 1]) fail("spectrum wavelengths must increase");
            g = std::gcd(g, milli[k] - milli[k - 1]);
        }
    }
    const long long count = (milli.back() - milli.front()) / g + 1;
    if (count > 100000) fail("spectrum wavelengths too finely spaced to tabulate");
    Spectrum s;
    s.kind = kRegular;
    s.type = type;
    s.lo = lambdas.front();
    s.hi = lambdas.back();
    s.values.resize((size_t)count);
    for (long long k = 0; k < count; ++k)
        s.values[(size_t)k] = irregularAt(lambdas, values, (float)((double)(milli.front() + k * g) / 1000.0));
    return s;
}

bool spectrumType(const std::string& s, int* type) {
    if (s == "Reflectance") *type = kReflectance;
    else if (s == "LightSource" || s == "Illuminant") *type = kLightSource;
    else if (s == "IoR" || s == "IndexOfRefraction") *type = kIndexOfRefraction;
    else return false;
    return true;
}

// ------------------------------------------------------------------ lexer

enum class Tok { End, Int, Real, Bool, String, Id, Op, Kw };

struct Token {
    Tok kind = Tok::End;
    std::string text;
    int32_t i = 0;
    double r = 0;
    bool b = false;
    int line = 1;
};

std::vector<Token> lex(const std::string& src) {
    std::vector<Token> out;
    size_t p = 0;
    int line = 1;
    const size_t n = src.size();
    auto isDigit = [](char c) { return c >= '0' && c <= '9'; };
    auto isAlpha = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; };
    while (p < n) {
        const char c = src[p];
        if (c == '\n') { ++line; ++p; continue; }
        if (c == ' ' || c == '\t' || c == '\r') { ++p; continue; }
        if (c == '/' && p + 1 < n && src[p + 1] == '/') {
            while (p < n && src[p] != '\n') ++p;
            continue;
        }
        if (c == '/' && p + 1 < n && src[p + 1] == '*') {
            p += 2;
            while (p + 1 < n && !(src[p] == '*' && src[p + 1] == '/')) { if (src[p] == '\n') ++line; ++p; }
            p = std::min(n, p + 2);
            continue;
        }
        Token t;
        t.line = line;
        if (c == '"') {
            size_t q = p + 1;
            std::string s;
            while (q < n && src[q] != '"') {
                if (src[q] == '\n') fail("line " + std::to_string(line) + ": unterminated string");
                if (src[q] == '\\' && q + 1 < n) { s += src[q + 1]; q += 2; continue; }
                s += src[q++];
            }
            if (q >= n) fail("line " + std::to_string(line) + ": unterminated string");
            t.kind = Tok::String;
            t.text = s;
            p = q + 1;
            out.push_back(t);
            continue;
        }
        if (isDigit(c) || (c == '.' && p + 1 < n && isDigit(src[p + 1]))) {
            // REALNUMBER {FLOAT}{EXPONENT}? against INTEGER {NONZERO}{DIGIT}*|0:
            // the longer match, INTEGER on a tie
            size_t q = p;
            bool real = false;
            if (src[q] == '0' && !(q + 1 < n && (src[q + 1] == '.'))) {
                ++q;   // "0" alone
            } else {
                while (q < n && isDigit(src[q])) ++q;
                if (q < n && src[q] == '.') { real = true; ++q; while (q < n && isDigit(src[q])) ++q; }
            }
            if (q < n && (src[q] == 'e' || src[q] == 'E')) {
                size_t e = q + 1;
                if (e < n && (src[e] == '+' || src[e] == '-')) ++e;
                if (e < n && isDigit(src[e])) {
                    while (e < n && isDigit(src[e])) ++e;
                    q = e;
                    real = true;
                }
            }
            const std::string num = src.substr(p, q - p);
            if (real) { t.kind = Tok::Real; t.r = std::atof(num.c_str()); }
            else { t.kind = Tok::Int; t.i = (int32_t)std::atoi(num.c_str()); }
            t.text = num;
            p = q;
            out.push_back(t);
            continue;
        }
        if (isAlpha(c)) {
            size_t q = p;
            while (q < n && (isAlpha(src[q]) || isDigit(src[q]))) ++q;
            const std::string id = src.substr(p, q - p);
            if (id == "True" || id == "true" || id == "False" || id == "false") {
                t.kind = Tok::Bool;
                t.b = id == "True" || id == "true";
            } else if (id == "if" || id == "else" || id == "for" || id == "function" || id == "return") {
                t.kind = Tok::Kw;
            } else {
                t.kind = Tok::Id;
            }
            t.text = id;
            p = q;
            out.push_back(t);
            continue;
        }
        static const char* ops[] = {"<=", ">=", "==", "!=", "&&", "||", "+=", "-=", "*=", "/=", "%=", "++", "--",
                                    "<", ">", "=", "+", "-", "*", "/", "%", "!", ",", ":", ";", "(", ")", "{", "}",
                                    "[", "]"};
        bool matched = false;
        for (const char* op : ops) {
            const size_t len = strlen(op);
            if (src.compare(p, len, op) == 0) {
                t.kind = Tok::Op;
                t.text = op;
                p += len;
                matched = true;
                break;
            }
        }
        if (!matched) fail("line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'");
        out.push_back(t);
    }
    Token e;
    e.line = line;
    out.push_back(e);
    return out;
}

// ------------------------------------------------------------------ interpreter

struct Interp {
    std::vector<Token> toks;
    size_t pos = 0;
    std::map<std::string, Value> vars;
    std::map<std::string, Builtin> builtins;
    SceneFile* out = nullptr;

    const Token& peek(size_t k = 0) const { return toks[std::min(pos + k, toks.size() - 1)]; }
    bool isOp(const char* op, size_t k = 0) const { return peek(k).kind == Tok::Op && peek(k).text == op; }
    [[noreturn]] void error(const std::string& m) const { fail("line " + std::to_string(peek().line) + ": " + m); }
    void expect(const char* op) {
        if (!isOp(op)) error(std::string("expected '") + op + "'");
        ++pos;
    }

    void run() {
        while (peek().kind != Tok::End) statement();
    }

    void statement() {
        if (peek().kind == Tok::Kw) error("'" + peek().text + "' statements are outside the supported subset");
        if (isOp("{")) {
            ++pos;
            while (!isOp("}")) {
                if (peek().kind == Tok::End) error("missing '}'");
                statement();
            }
            ++pos;
            return;
        }
        expression();
        expect(";");
    }

    // Expression: assignment, then || && == != < > <= >= + -, Term: * / %,
    // unary + - !, SingleTerm with [] indexing (SceneParser.yy precedences)
    Value expression() {
        if (peek().kind == Tok::Id && peek(1).kind == Tok::Op) {
            const std::string& op = peek(1).text;
            if (op == "=" || op == "+=" || op == "-=" || op == "*=" || op == "/=" || op == "%=") {
                const std::string name = peek().text;
                pos += 2;
                Value rhs = expression();
                if (op != "=") {
                    auto it = vars.find(name);
                    if (it == vars.end()) error("undefined variable '" + name + "'");
                    rhs = binary(it->second, op.substr(0, 1), rhs);
                }
                vars[name] = rhs;
                return rhs;
            }
        }
        return logicOr();
    }
    Value logicOr() {
        Value v = logicAnd();
        while (isOp("||")) { ++pos; v = binary(v, "||", logicAnd()); }
        return v;
    }
    Value logicAnd() {
        Value v = equality();
        while (isOp("&&")) { ++pos; v = binary(v, "&&", equality()); }
        return v;
    }
    Value equality() {
        Value v = relation();
        while (isOp("==") || isOp("!=")) { const std::string op = peek().text; ++pos; v = binary(v, op, relation()); }
        return v;
    }
    Value relation() {
        Value v = additive();
        while (isOp("<") || isOp(">") || isOp("<=") || isOp(">=")) {
            const std::string op = peek().text;
            ++pos;
            v = binary(v, op, additive());
        }
        return v;
    }
    Value additive() {
        Value v = term();
        while (isOp("+") || isOp("-")) { const std::string op = peek().text; ++pos; v = binary(v, op, term()); }
        return v;
    }
    Value term() {
        Value v = unary();
        while (isOp("*") || isOp("/") || isOp("%")) { const std::string op = peek().text; ++pos; v = binary(v, op, unary()); }
        return v;
    }
    Value unary() {
        if (isOp("-") || isOp("+") || isOp("!")) {
            const std::string op = peek().text;
            ++pos;
            Value v = single();
            if (op == "-") {
                if (v.ty == Ty::Integer) return mkInt(-v.i);
                if (v.ty == Ty::Real) return mkReal(-v.r);
                error("unary - of a " + std::string(tyName(v.ty)));
            }
            if (op == "!") { Value x; x.ty = Ty::Bool; x.b = !truth(v); return x; }
            return v;
        }
        if (isOp("++") || isOp("--")) error("increment operators are outside the supported subset");
        return single();
    }
    Value single() {
        Value v = primary();
        while (isOp("[")) {
            ++pos;
            Value k = expression();
            expect("]");
            if (v.ty != Ty::Tuple) error("indexing a " + std::string(tyName(v.ty)));
            if (k.ty == Ty::Integer) {
                if (k.i < 0 || (size_t)k.i >= v.tup->unnamed.size()) error("tuple index out of range");
                v = v.tup->unnamed[(size_t)k.i];
            } else if (k.ty == Ty::String) {
                const Value* f = nullptr;
                for (auto& kv : v.tup->named)
                    if (kv.first == k.s) f = &kv.second;
                if (!f) error("no tuple element '" + k.s + "'");
                v = *f;
            } else {
                error("bad tuple index");
            }
        }
        return v;
    }
    // Parameter: Expression or Expression ":" Expression (a named element)
    void parameter(TupleV* t) {
        Value a = expression();
        if (isOp(":")) {
            ++pos;
            if (a.ty != Ty::String) error("an element name must be a string");
            Value b = expression();
            t->named.emplace_back(a.s, b);
        } else {
            t->unnamed.push_back(a);
        }
    }
    Value primary() {
        const Token& t = peek();
        switch (t.kind) {
            case Tok::Int: ++pos; return mkInt(t.i);
            case Tok::Real: ++pos; return mkReal(t.r);
            case Tok::Bool: { ++pos; Value x; x.ty = Ty::Bool; x.b = t.b; return x; }
            case Tok::String: ++pos; return mkStr(t.text);
            case Tok::Id: {
                const std::string name = t.text;
                ++pos;
                if (isOp("(")) {
                    ++pos;
                    Value args = mkTuple();
                    if (!isOp(")")) {
                        parameter(args.tup.get());
                        while (isOp(",")) { ++pos; parameter(args.tup.get()); }
                    }
                    expect(")");
                    return call(name, *args.tup);
                }
                auto it = vars.find(name);
                if (it == vars.end()) error("undefined variable '" + name + "'");
                return it->second;
            }
            case Tok::Op:
                if (t.text == "(") {
                    ++pos;
                    Value tup = mkTuple();
                    if (isOp(",")) {   // "(" "," ")": the empty tuple
                        ++pos;
                        expect(")");
                        return tup;
                    }
                    // "(" Expression ")" unless a "," or ":" follows the first
                    // element (each element is evaluated once)
                    parameter(tup.tup.get());
                    if (isOp(")") && tup.tup->named.empty()) {
                        ++pos;
                        return tup.tup->unnamed[0];
                    }
                    while (isOp(",")) {
                        ++pos;
                        if (isOp(")")) break;   // "(" Parameter "," ")"
                        parameter(tup.tup.get());
                    }
                    expect(")");
                    return tup;
                }
                break;
            default:
                break;
        }
        error("unexpected '" + t.text + "'");
    }

    static bool truth(const Value& v) {
        if (v.ty == Ty::Bool) return v.b;
        if (v.ty == Ty::Integer) return v.i != 0;
        if (v.ty == Ty::Real) return v.r != 0;
        return false;
    }
    Value binary(const Value& a, const std::string& op, const Value& b) {
        const bool num = (a.ty == Ty::Integer || a.ty == Ty::Real || a.ty == Ty::Bool) &&
                         (b.ty == Ty::Integer || b.ty == Ty::Real || b.ty == Ty::Bool);
        if (op == "&&" || op == "||") {
            Value x;
            x.ty = Ty::Bool;
            x.b = op == "&&" ? truth(a) && truth(b) : truth(a) || truth(b);
            return x;
        }
        if (num) {
            const bool ints = a.ty != Ty::Real && b.ty != Ty::Real;
            if (op == "<" || op == ">" || op == "<=" || op == ">=" || op == "==" || op == "!=") {
                const double x = a.real(), y = b.real();
                Value r;
                r.ty = Ty::Bool;
                r.b = op == "<" ? x < y : op == ">" ? x > y : op == "<=" ? x <= y : op == ">=" ? x >= y
                                                                     : op == "==" ? x == y : x != y;
                return r;
            }
            if (ints) {
                const int32_t x = a.ty == Ty::Bool ? a.b : a.i, y = b.ty == Ty::Bool ? b.b : b.i;
                if (op == "+") return mkInt(x + y);
                if (op == "-") return mkInt(x - y);
                if (op == "*") return mkInt(x * y);
                if ((op == "/" || op == "%") && y == 0) error("integer division by zero");
                if (op == "/") return mkInt(x / y);
                if (op == "%") return mkInt(x % y);
            } else {
                const double x = a.real(), y = b.real();
                if (op == "+") return mkReal(x + y);
                if (op == "-") return mkReal(x - y);
                if (op == "*") return mkReal(x * y);
                if (op == "/") return mkReal(x / y);
                if (op == "%") return mkReal(std::fmod(x, y));
            }
        }
        if (op == "+" && a.ty == Ty::String && b.ty == Ty::String) return mkStr(a.s + b.s);
        if (op == "*" && a.ty == Ty::Matrix && b.ty == Ty::Matrix) {
            Value r;
            r.ty = Ty::Matrix;
            r.mat = mul(a.mat, b.mat);
            return r;
        }
        error("operator " + op + " on a " + tyName(a.ty) + " and a " + tyName(b.ty));
    }

    Value call(const std::string& name, const TupleV& params) {
        auto it = builtins.find(name);
        if (it == builtins.end()) {
            auto v = vars.find(name);
            if (v != vars.end()) error("calling '" + name + "': user functions are outside the supported subset");
            error("'" + name + "' is outside the supported subset");
        }
        const Builtin& f = it->second;
        for (size_t k = 0; k < f.sigs.size(); ++k) {
            Args args;
            if (mapArgs(params, f.sigs[k], &args)) return f.procs[k](*this, args);
        }
        error("the arguments do not match " + name + "()");
    }
};

// ------------------------------------------------------------------ builtins

std::shared_ptr<Obj> objOf(const Value& v) { return v.obj ? v.obj : std::make_shared<Obj>(); }

int addSpectrum(Interp& in, Spectrum s) {
    in.out->spectra.push_back(std::move(s));
    return (int)in.out->spectra.size() - 1;
}

Value spectrumValue(int index) {
    auto o = std::make_shared<Obj>();
    o->index = index;
    return mkObj(Ty::Spectrum, o);
}

std::vector<float> realsOf(Interp& in, const Value& tup, const char* what) {
    std::vector<float> v;
    for (const Value& e : tup.tup->unnamed) {
        if (!convertible(e.ty, Ty::Real)) in.error(std::string(what) + ": not a number");
        v.push_back((float)e.real());
    }
    return v;
}

Mat4 rotation(float angle, float ax, float ay, float az) {
    // Matrix4x4 rotate (libSLR/BasicTypes/Matrix4x4.cpp:111-133) with the
    // axis normalised as Vector3D normalize does
    const float len = std::sqrt(ax * ax + ay * ay + az * az);
    const float x = ax / len, y = ay / len, z = az / len;
    const float c = std::cos(angle), s = std::sin(angle);
    const float omc = 1 - c;
    Mat4 r = Mat4::identity();
    r.m[0][0] = x * x * omc + c;
    r.m[0][1] = x * y * omc + z * s;
    r.m[0][2] = z * x * omc - y * s;
    r.m[1][0] = x * y * omc - z * s;
    r.m[1][1] = y * y * omc + c;
    r.m[1][2] = y * z * omc + x * s;
    r.m[2][0] = z * x * omc + y * s;
    r.m[2][1] = y * z * omc - x * s;
    r.m[2][2] = z * z * omc + c;
    return r;
}

Value matrixValue(const Mat4& m) {
    Value v;
    v.ty = Ty::Matrix;
    v.mat = m;
    return v;
}

void installBuiltins(Interp& in) {
    auto& B = in.builtins;
    const Value R = mkStr("Reflectance");
    // Spectrum (API.cpp:336-470)
    B["Spectrum"].sigs = {
        {req("type", Ty::String), req("value", Ty::Real)},
        {opt("type", Ty::String, R), opt("space", Ty::String, mkStr("sRGB")), req("e0", Ty::Real), req("e1", Ty::Real),
         req("e2", Ty::Real)},
        {opt("type", Ty::String, R), req("minWL", Ty::Real), req("maxWL", Ty::Real), req("values", Ty::Tuple)},
        {opt("type", Ty::String, R), req("wls", Ty::Tuple), req("values", Ty::Tuple)},
        {req("type", Ty::String), req("name", Ty::String), opt("idx", Ty::Integer, mkInt(0))}};
    auto rgbSpectrum = [](Interp& in, const std::string& typeStr, const std::string& space, float e0, float e1, float e2) {
        Spectrum s;
        if (!spectrumType(typeStr, &s.type)) in.error("invalid spectrum type '" + typeStr + "'");
        if (space != "sRGB") in.error("colour space '" + space + "' is outside the supported subset (sRGB)");
        if (s.type == kIndexOfRefraction) in.error("RGB indices of refraction are outside the supported subset");
        s.kind = kRgb;
        s.rgb[0] = e0; s.rgb[1] = e1; s.rgb[2] = e2;
        return spectrumValue(addSpectrum(in, s));
    };
    B["Spectrum"].procs = {
        [rgbSpectrum](Interp& in, const Args& a) {
            const float v = (float)a.at("value").r;
            return rgbSpectrum(in, a.at("type").s, "sRGB", v, v, v);
        },
        [rgbSpectrum](Interp& in, const Args& a) {
            return rgbSpectrum(in, a.at("type").s, a.at("space").s, (float)a.at("e0").r, (float)a.at("e1").r,
                               (float)a.at("e2").r);
        },
        [](Interp& in, const Args& a) {
            Spectrum s;
            if (!spectrumType(a.at("type").s, &s.type)) in.error("invalid spectrum type");
            s.kind = kRegular;
            s.lo = (float)a.at("minWL").r;
            s.hi = (float)a.at("maxWL").r;
            s.values = realsOf(in, a.at("values"), "Spectrum values");
            if (s.values.empty()) in.error("empty spectrum");
            return spectrumValue(addSpectrum(in, s));
        },
        [](Interp& in, const Args& a) {
            int type = 0;
            if (!spectrumType(a.at("type").s, &type)) in.error("invalid spectrum type");
            std::vector<float> wls = realsOf(in, a.at("wls"), "Spectrum wavelengths");
            std::vector<float> vals = realsOf(in, a.at("values"), "Spectrum values");
            if (wls.size() != vals.size() || wls.empty()) in.error("the sizes of the wavelengths and the values differ");
            return spectrumValue(addSpectrum(in, irregularAsRegular(type, wls, vals)));
        },
        [](Interp& in, const Args& a) {
            const std::string type = a.at("type").s, name = a.at("name").s;
            const int32_t idx = a.at("idx").i;
            const tables::LibrarySpectrum* found = nullptr;
            if (type == "Illuminant")
                for (const auto& e : tables::kIlluminants)
                    if (name == e.name && idx == 0) found = &e;
            if (!found) in.error("library spectrum " + type + " '" + name + "' is outside the supported subset");
            Spectrum s;
            s.kind = kRegular;
            s.type = kLightSource;
            s.lo = found->minLambda;
            s.hi = found->maxLambda;
            s.values.assign(found->values, found->values + found->numSamples);
            return spectrumValue(addSpectrum(in, s));
        }};
    // SpectrumTexture / FloatTexture (builtin_texture.cpp:70-82, 230-242): constants
    B["SpectrumTexture"].sigs = {{req("spectrum", Ty::Spectrum)}};
    B["SpectrumTexture"].procs = {[](Interp&, const Args& a) {
        auto o = std::make_shared<Obj>();
        o->spectrum = a.at("spectrum").obj->index;
        return mkObj(Ty::SpectrumTexture, o);
    }};
    B["FloatTexture"].sigs = {{req("value", Ty::Real)}};
    B["FloatTexture"].procs = {[](Interp&, const Args& a) {
        auto o = std::make_shared<Obj>();
        o->value = (float)a.at("value").r;
        return mkObj(Ty::FloatTexture, o);
    }};
    // createSurfaceMaterial (API.cpp:520-728)
    B["createSurfaceMaterial"].sigs = {{req("type", Ty::String), req("params", Ty::Tuple)}};
    B["createSurfaceMaterial"].procs = {[](Interp& in, const Args& a) {
        const std::string type = a.at("type").s;
        const TupleV& params = *a.at("params").tup;
        Material m;
        Args p;
        auto configure = [&](const std::vector<Arg>& sig) {
            if (!mapArgs(params, sig, &p)) in.error("the parameters do not match material '" + type + "'");
        };
        if (type == "matte") {
            configure({req("reflectance", Ty::SpectrumTexture), optNull("sigma", Ty::FloatTexture)});
            m.kind = kMatte;
            m.spec0 = p.at("reflectance").obj->spectrum;
            if (p.at("sigma").ty == Ty::FloatTexture) {
                m.hasSigma = true;
                m.sigma = p.at("sigma").obj->value;
            }
        } else if (type == "microfacet metal") {
            configure({req("eta", Ty::SpectrumTexture), req("k", Ty::SpectrumTexture), req("alpha_gx", Ty::FloatTexture),
                       req("alpha_gy", Ty::FloatTexture)});
            m.kind = kMicrofacetMetal;
            m.spec0 = p.at("eta").obj->spectrum;
            m.spec1 = p.at("k").obj->spectrum;
            m.alphaX = p.at("alpha_gx").obj->value;
            m.alphaY = p.at("alpha_gy").obj->value;
        } else if (type == "microfacet glass") {
            configure({req("etaExt", Ty::SpectrumTexture), req("etaInt", Ty::SpectrumTexture),
                       req("alpha_gx", Ty::FloatTexture), req("alpha_gy", Ty::FloatTexture)});
            m.kind = kMicrofacetGlass;
            m.spec0 = p.at("etaExt").obj->spectrum;
            m.spec1 = p.at("etaInt").obj->spectrum;
            m.alphaX = p.at("alpha_gx").obj->value;
            m.alphaY = p.at("alpha_gy").obj->value;
        } else if (type == "emitter") {
            configure({req("scatter", Ty::SurfaceMaterial), req("emitter", Ty::EmitterProperty)});
            m.kind = kEmitter;
            m.base = p.at("scatter").obj->index;
            m.emittance = p.at("emitter").obj->spectrum;
        } else {
            in.error("material type '" + type + "' is outside the supported subset");
        }
        in.out->materials.push_back(m);
        auto o = std::make_shared<Obj>();
        o->index = (int)in.out->materials.size() - 1;
        return mkObj(Ty::SurfaceMaterial, o);
    }};
    B["createEmitterSurfaceProperty"].sigs = {{req("type", Ty::String), req("params", Ty::Tuple)}};
    B["createEmitterSurfaceProperty"].procs = {[](Interp& in, const Args& a) {
        if (a.at("type").s != "diffuse") in.error("emitter type '" + a.at("type").s + "' is outside the supported subset");
        Args p;
        if (!mapArgs(*a.at("params").tup, {req("emittance", Ty::SpectrumTexture)}, &p))
            in.error("the parameters do not match emitter 'diffuse'");
        auto o = std::make_shared<Obj>();
        o->spectrum = p.at("emittance").obj->spectrum;
        return mkObj(Ty::EmitterProperty, o);
    }};
    // nodes and transforms (API.cpp:1020-1073, builtin_transform.cpp:15-71)
    B["createNode"].sigs = {{}};
    B["createNode"].procs = {[](Interp&, const Args&) {
        auto o = std::make_shared<Obj>();
        o->node = std::make_shared<Node>();
        return mkObj(Ty::InternalNode, o);
    }};
    B["addChild"].sigs = {{req("parent", Ty::InternalNode), req("child", Ty::Node)}};
    B["addChild"].procs = {[](Interp&, const Args& a) {
        a.at("parent").obj->node->children.push_back(a.at("child").obj->node);
        return Value();
    }};
    B["setTransform"].sigs = {{req("node", Ty::InternalNode), req("transform", Ty::Transform)}};
    B["setTransform"].procs = {[](Interp&, const Args& a) {
        a.at("node").obj->node->local = a.at("transform").mat;
        return Value();
    }};
    B["translate"].sigs = {{req("x", Ty::Real), req("y", Ty::Real), req("z", Ty::Real)}};
    B["translate"].procs = {[](Interp&, const Args& a) {
        Mat4 m = Mat4::identity();
        m.m[3][0] = (float)a.at("x").r;
        m.m[3][1] = (float)a.at("y").r;
        m.m[3][2] = (float)a.at("z").r;
        return matrixValue(m);
    }};
    B["scale"].sigs = {{req("s", Ty::Real)}, {req("x", Ty::Real), req("y", Ty::Real), req("z", Ty::Real)}};
    auto scaleMat = [](float x, float y, float z) {
        Mat4 m = Mat4::identity();
        m.m[0][0] = x;
        m.m[1][1] = y;
        m.m[2][2] = z;
        return matrixValue(m);
    };
    B["scale"].procs = {[scaleMat](Interp&, const Args& a) {
                            const float s = (float)a.at("s").r;
                            return scaleMat(s, s, s);
                        },
                        [scaleMat](Interp&, const Args& a) {
                            return scaleMat((float)a.at("x").r, (float)a.at("y").r, (float)a.at("z").r);
                        }};
    B["rotate"].sigs = {{req("angle", Ty::Real), req("axis", Ty::Tuple)}};
    B["rotate"].procs = {[](Interp& in, const Args& a) {
        std::vector<float> ax = realsOf(in, a.at("axis"), "rotate axis");
        if (ax.size() != 3) in.error("rotate: the axis needs three components");
        return matrixValue(rotation((float)a.at("angle").r, ax[0], ax[1], ax[2]));
    }};
    for (int k = 0; k < 3; ++k) {
        const char* names[3] = {"rotateX", "rotateY", "rotateZ"};
        B[names[k]].sigs = {{req("angle", Ty::Real)}};
        B[names[k]].procs = {[k](Interp&, const Args& a) {
            return matrixValue(rotation((float)a.at("angle").r, k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f,
                                        k == 2 ? 1.0f : 0.0f));
        }};
    }
    // createVertex (API.cpp:310-334): the position is what the mesh keeps
    // (normals and tangents are recomputed by createMesh)
    B["createVertex"].sigs = {
        {req("position", Ty::Tuple), req("normal", Ty::Tuple), req("tangent", Ty::Tuple), req("texCoord", Ty::Tuple)}};
    B["createVertex"].procs = {[](Interp& in, const Args& a) {
        std::vector<float> p = realsOf(in, a.at("position"), "vertex position");
        if (p.size() != 3) in.error("a vertex position needs three components");
        auto o = std::make_shared<Obj>();
        o->vtx[0] = p[0]; o->vtx[1] = p[1]; o->vtx[2] = p[2];
        return mkObj(Ty::Vertex, o);
    }};
    // createMesh (API.cpp:795-870)
    B["createMesh"].sigs = {
        {req("vertices", Ty::Tuple), req("matGroups", Ty::Tuple), opt("axisForRadialTangent", Ty::Integer, mkInt(-1))}};
    B["createMesh"].procs = {[](Interp& in, const Args& a) {
        auto node = std::make_shared<Node>();
        node->kind = Node::kMeshNode;
        for (const Value& v : a.at("vertices").tup->unnamed) {
            Value vtx = v;
            if (v.ty == Ty::Tuple) vtx = in.call("createVertex", *v.tup);
            if (vtx.ty != Ty::Vertex) in.error("a mesh vertex must be a vertex or a tuple");
            node->pos.insert(node->pos.end(), vtx.obj->vtx, vtx.obj->vtx + 3);
        }
        const uint32_t nv = (uint32_t)(node->pos.size() / 3);
        for (const Value& g : a.at("matGroups").tup->unnamed) {
            if (g.ty != Ty::Tuple) in.error("a material group must be a tuple");
            Args p;
            if (!mapArgs(*g.tup, {req("mat", Ty::SurfaceMaterial), optNull("normal", Ty::NormalTexture),
                                  optNull("alpha", Ty::FloatTexture), req("triangles", Ty::Tuple)}, &p))
                in.error("the elements do not match a material group");
            if (p.at("normal").ty != Ty::Void || p.at("alpha").ty != Ty::Void)
                in.error("normal and alpha maps are outside the supported subset");
            const uint32_t mat = (uint32_t)p.at("mat").obj->index;
            for (const Value& t : p.at("triangles").tup->unnamed) {
                Args tri;
                if (t.ty != Ty::Tuple ||
                    !mapArgs(*t.tup, {req("v0", Ty::Integer), req("v1", Ty::Integer), req("v2", Ty::Integer)}, &tri))
                    in.error("a triangle must be three vertex indices");
                for (const char* k : {"v0", "v1", "v2"}) {
                    const int32_t vi = tri.at(k).i;
                    if (vi < 0 || (uint32_t)vi >= nv) in.error("triangle vertex index out of range");
                    node->idx.push_back((uint32_t)vi);
                }
                node->mat.push_back(mat);
            }
        }
        if (a.at("axisForRadialTangent").i != -1) in.error("radial tangents are outside the supported subset");
        auto o = std::make_shared<Obj>();
        o->node = node;
        return mkObj(Ty::SurfaceNode, o);
    }};
    // createPerspectiveCamera (API.cpp:1176-1197)
    B["createPerspectiveCamera"].sigs = {
        {opt("sensitivity", Ty::Real, mkReal(0.0)), opt("aspect", Ty::Real, mkReal(1.0)),
         opt("fovY", Ty::Real, mkReal(0.5235987756)), opt("radius", Ty::Real, mkReal(0.0)),
         opt("imgDist", Ty::Real, mkReal(0.02)), opt("objDist", Ty::Real, mkReal(5.0))}};
    B["createPerspectiveCamera"].procs = {[](Interp&, const Args& a) {
        auto node = std::make_shared<Node>();
        node->kind = Node::kCameraNode;
        node->cam.sensitivity = (float)a.at("sensitivity").r;
        node->cam.aspect = (float)a.at("aspect").r;
        node->cam.fovY = (float)a.at("fovY").r;
        node->cam.radius = (float)a.at("radius").r;
        node->cam.imgDist = (float)a.at("imgDist").r;
        node->cam.objDist = (float)a.at("objDist").r;
        auto o = std::make_shared<Obj>();
        o->node = node;
        return mkObj(Ty::CameraNode, o);
    }};
    // setRenderer / setRenderSettings (API.cpp:1214-1327)
    B["setRenderer"].sigs = {{req("method", Ty::String), opt("config", Ty::Tuple, mkTuple())}};
    B["setRenderer"].procs = {[](Interp& in, const Args& a) {
        const std::string method = a.at("method").s;
        if (method != "PT" && method != "Volumetric PT" && method != "BPT")
            in.error("renderer '" + method + "' is outside the supported subset (PT, Volumetric PT, BPT)");
        Args p;
        if (!mapArgs(*a.at("config").tup, {opt("samples", Ty::Integer, mkInt(8))}, &p))
            in.error("the renderer configuration does not match");
        in.out->method = method;
        in.out->spp = (uint32_t)p.at("samples").i;
        in.out->hasRenderer = true;
        return Value();
    }};
    B["setRenderSettings"].sigs = {{opt("numThreads", Ty::Integer, mkInt(0)), opt("width", Ty::Integer, mkInt(1024)),
                                    opt("height", Ty::Integer, mkInt(1024)), opt("timeStart", Ty::Real, mkReal(0.0)),
                                    opt("timeEnd", Ty::Real, mkReal(0.0)), opt("brightness", Ty::Real, mkReal(1.0)),
                                    opt("rngSeed", Ty::Integer, mkInt(1509761209))}};
    B["setRenderSettings"].procs = {[](Interp& in, const Args& a) {
        if (a.at("width").i <= 0 || a.at("height").i <= 0) in.error("image size must be positive");
        in.out->width = (uint32_t)a.at("width").i;
        in.out->height = (uint32_t)a.at("height").i;
        in.out->brightness = (float)a.at("brightness").r;
        in.out->rngSeed = (uint32_t)a.at("rngSeed").i;
        return Value();
    }};
}

// ------------------------------------------------------------------ assembly

float len3(const float v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// normalize(v) of a Vector3D or Normal3D: v / length, which those types
// compute as a multiply by the float reciprocal (Vector3D.h:31,115-118,
// Normal3D.h:36,108-111).
void normalize3(float v[3]) {
    const float r = 1.0f / len3(v);
    v[0] = v[0] * r;
    v[1] = v[1] * r;
    v[2] = v[2] * r;
}
// Normal3D::normalize(), the member: its reciprocal length in double,
// rounded to float (Normal3D.h:53-58).
void normalizeMember3(float v[3]) {
    const float r = (float)(1.0 / (double)len3(v));
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
}

void emitMesh(const Node& node, const Mat4& world, const float inv[4][4], bool identity, SceneFile* out) {
    const size_t nv = node.pos.size() / 3;
    std::vector<float> nrm(3 * nv, 0.0f);
    // face normals summed per vertex in material-group order
    // (recalculateNormalAndTangents), from the object-space positions
    for (size_t t = 0; t < node.idx.size() / 3; ++t) {
        const float* p0 = &node.pos[3 * node.idx[3 * t]];
        const float* p1 = &node.pos[3 * node.idx[3 * t + 1]];
        const float* p2 = &node.pos[3 * node.idx[3 * t + 2]];
        const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float f[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        normalize3(f);
        for (int k = 0; k < 3; ++k) {
            float* n = &nrm[3 * node.idx[3 * t + k]];
            n[0] = n[0] + f[0];
            n[1] = n[1] + f[1];
            n[2] = n[2] + f[2];
        }
    }
    Mesh m;
    m.pos.resize(3 * nv);
    m.nrm.resize(3 * nv);
    m.tan.resize(3 * nv);
    for (size_t v = 0; v < nv; ++v) {
        float n[3] = {nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]};
        normalizeMember3(n);
        float t[3];
        coordinateTangent(n, t);
        const float* p = &node.pos[3 * v];
        float wp[3] = {p[0], p[1], p[2]}, wn[3] = {n[0], n[1], n[2]}, wt[3] = {t[0], t[1], t[2]};
        if (!identity) {
            // Matrix4x4 * Point3D, then the divide by w when it is not one
            float h[4];
            for (int r = 0; r < 4; ++r)
                h[r] = world.at(r, 0) * p[0] + world.at(r, 1) * p[1] + world.at(r, 2) * p[2] + world.at(r, 3) * 1.0f;
            if (h[3] != 1.0f)
                for (int r = 0; r < 3; ++r) h[r] = h[r] / h[3];
            wp[0] = h[0]; wp[1] = h[1]; wp[2] = h[2];
            // StaticTransform * Normal3D: the inverse's transpose (transform.h:44-49)
            wn[0] = inv[0][0] * n[0] + inv[0][1] * n[1] + inv[0][2] * n[2];
            wn[1] = inv[1][0] * n[0] + inv[1][1] * n[1] + inv[1][2] * n[2];
            wn[2] = inv[2][0] * n[0] + inv[2][1] * n[1] + inv[2][2] * n[2];
            normalize3(wn);
            // StaticTransform * Vector3D: the rows of the matrix
            for (int r = 0; r < 3; ++r) wt[r] = world.at(r, 0) * t[0] + world.at(r, 1) * t[1] + world.at(r, 2) * t[2];
            normalize3(wt);
        }
        for (int k = 0; k < 3; ++k) {
            m.pos[3 * v + k] = wp[k];
            m.nrm[3 * v + k] = wn[k];
            m.tan[3 * v + k] = wt[k];
        }
    }
    m.idx = node.idx;
    m.mat = node.mat;
    out->meshes.push_back(std::move(m));
}

bool isIdentity(const Mat4& m) {
    const Mat4 I = Mat4::identity();
    return memcmp(m.m, I.m, sizeof(m.m)) == 0;
}

void assemble(const Node& node, const Mat4& parentWorld, bool root, SceneFile* out, int depth) {
    if (depth > 256) fail("node graph too deep (a cycle?)");
    const Mat4 world = root ? node.local : mul(parentWorld, node.local);
    for (const auto& child : node.children) {
        if (child->kind == Node::kInternalNode) {
            assemble(*child, world, false, out, depth + 1);
        } else if (child->kind == Node::kMeshNode) {
            float inv[4][4];
            const bool ident = isIdentity(world);
            if (!ident && !invertMatrix4(world.m, inv)) fail("singular node transform");
            emitMesh(*child, world, inv, ident, out);
        } else {
            if (out->camera.present) fail("more than one camera in the scene");
            out->camera = child->cam;
            out->camera.present = true;
            memcpy(out->camera.l2w, world.m, sizeof(out->camera.l2w));
        }
    }
}

}  // namespace

bool readSceneText(const std::string& text, SceneFile* out, std::string* err) {
    *out = SceneFile();
    try {
        Interp in;
        in.out = out;
        in.toks = lex(text);
        installBuiltins(in);
        auto rootObj = std::make_shared<Obj>();
        rootObj->node = std::make_shared<Node>();
        in.vars["root"] = mkObj(Ty::InternalNode, rootObj);
        in.run();
        assemble(*rootObj->node, Mat4::identity(), true, out, 0);
        if (!out->camera.present) fail("the scene has no camera");
        if (out->meshes.empty()) fail("the scene has no mesh");
        return true;
    } catch (const ScriptError& e) {
        if (err) *err = e.msg;
        return false;
    }
}

}  // namespace scenefile
}  // namespace slrh

// sceneio.cpp -- the reference's scene input/output on the host: the XML
// scene configuration (Config::load, config.cpp:432-575), the .obj / .shp
// shape loaders (Mesh::loadShape, mesh.cpp:5-130) and the PFM writer
// (Film::savePFM after Film::setBuffer, film.h:36-42,100-110).
//
// Only what feeds the GPU path is kept: the four hot object classes, static
// transforms, motion, the pinhole camera and the path tracer's settings;
// depth of field, textures, volumes and the other classes are refused with an
// error naming them.  World-space vertices and vertex normals are computed on
// the host with the reference's float operation order (matrix * vec4,
// transpose(inverse(M)) for normals), so the triangles handed to
// wr_scene_build are the ones the reference's Scene::buildKDTree would see.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/walkray.h"
#include "wr_sceneio.h"

namespace wr {
namespace io {

// ---- 4x4 matrices (nvMatrix.h semantics: element(row, col)) -----------------
Mat4 Mat4::identity() {
    Mat4 m;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m.e[i][j] = i == j ? 1.f : 0.f;
    return m;
}

// matrix * (x, y, z, w), each row summed left to right (nvMatrix.h:424-432)
void Mat4::apply(const float* v, float w, float* out) const {
    for (int i = 0; i < 3; i++) out[i] = v[0] * e[i][0] + v[1] * e[i][1] + v[2] * e[i][2] + w * e[i][3];
}

Mat4 transpose(const Mat4& m) {
    Mat4 t;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) t.e[i][j] = m.e[j][i];
    return t;
}

// Gauss-Jordan elimination with scaled partial pivoting in float, the
// algorithm of nv::inverse (nvMatrix.h:242-378): the same pivot choice and
// operation order, so the normal matrix is bit-identical.  A singular matrix
// gives the identity (the reference returns its default-constructed matrix).
Mat4 inverse(const Mat4& m) {
    float rows[4][8];
    float* s[4] = {rows[0], rows[1], rows[2], rows[3]};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            s[i][j] = m.e[i][j];
            s[i][j + 4] = i == j ? 1.f : 0.f;
        }
    float scale[4];
    for (int i = 0; i < 4; i++) {
        scale[i] = std::fabs(s[i][0]);
        for (int j = 1; j < 4; j++)
            if (std::fabs(s[i][j]) > scale[i]) scale[i] = std::fabs(s[i][j]);
        if (scale[i] == 0.f) return Mat4::identity();
    }
    for (int i = 0; i < 4; i++) {
        int piv = i;
        float best = std::fabs(s[i][i] / scale[i]);
        for (int p = i + 1; p < 4; p++) {
            const float c = std::fabs(s[p][i] / scale[p]);
            if (c > best) {
                best = c;
                piv = p;
            }
        }
        if (piv != i) {
            std::swap(s[i], s[piv]);
            std::swap(scale[i], scale[piv]);
        }
        for (int j = i + 1; j < 4; j++) {
            const float f = s[j][i] / s[i][i];
            s[j][i] = 0.f;
            for (int k = i + 1; k < 8; k++) {
                const float p = f * s[i][k];
                s[j][k] = s[j][k] - p;
            }
        }
    }
    if (s[3][3] == 0.f) return Mat4::identity();
    for (int i = 3; i > 0; i--)
        for (int j = i - 1; j > -1; j--) {
            const float f = s[j][i] / s[i][i];
            for (int k = j + 1; k < 8; k++) {
                const float p = f * s[i][k];
                s[j][k] = s[j][k] - p;
            }
        }
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.e[i][j] = s[i][j + 4] / s[i][i];
    return r;
}

// matrix * matrix, each element summed from 0 over c in order (nvMatrix.h:409-421)
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.f;
            for (int c = 0; c < 4; c++) acc += a.e[i][c] * b.e[c][j];
            r.e[i][j] = acc;
        }
    return r;
}

// ---- AnimatedTransform (motionBlur.h:6-34, motionBlur.cpp:6-87) with the
// Quaternion it uses (quaternion.h:9-75, quaternion.cpp:5-85) --------------
static float qdot(const Quat& a, const Quat& b) {  // Dot: v.dot(v) (from 0) + w * w
    float r = 0.f;
    r += a.v[0] * b.v[0];
    r += a.v[1] * b.v[1];
    r += a.v[2] * b.v[2];
    return r + a.w * b.w;
}
static Quat qmul(const Quat& q, float f) { return Quat{{q.v[0] * f, q.v[1] * f, q.v[2] * f}, q.w * f}; }
static Quat qadd(const Quat& a, const Quat& b) {
    return Quat{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}, a.w + b.w};
}
static Quat qsub(const Quat& a, const Quat& b) {
    return Quat{{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}, a.w - b.w};
}
// operator/(float): the vector part divides as vec3 (a zero divisor leaves it), w as a float
static Quat qdiv(const Quat& q, float f) {
    Quat r = q;
    if (f != 0.f)
        for (int k = 0; k < 3; k++) r.v[k] = q.v[k] / f;
    r.w = q.w / f;
    return r;
}
static Quat qnormalize(const Quat& q) { return qdiv(q, std::sqrt(qdot(q, q))); }

// Quaternion::ToTransform (quaternion.cpp:6-31): the rotation block of an identity matrix
static Mat4 quat_to_matrix(const Quat& q) {
    const float x = q.v[0], y = q.v[1], z = q.v[2], w = q.w;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float wx = x * w, wy = y * w, wz = z * w;
    Mat4 m = Mat4::identity();
    m.e[0][0] = 1.f - 2.f * (yy + zz);
    m.e[0][1] = 2.f * (xy + wz);
    m.e[0][2] = 2.f * (xz - wy);
    m.e[1][0] = 2.f * (xy - wz);
    m.e[1][1] = 1.f - 2.f * (xx + zz);
    m.e[1][2] = 2.f * (yz + wx);
    m.e[2][0] = 2.f * (xz + wy);
    m.e[2][1] = 2.f * (yz - wx);
    m.e[2][2] = 1.f - 2.f * (xx + yy);
    return m;
}

// Quaternion::Quaternion(const Transform&) (quaternion.cpp:34-70)
static Quat matrix_to_quat(const Mat4& m) {
    Quat q{{0.f, 0.f, 0.f}, 1.f};
    const float trace = m.e[0][0] + m.e[1][1] + m.e[2][2];
    if (trace > 0.f) {
        float s = std::sqrt((float)((double)trace + 1.0));
        q.w = s / 2.0f;
        s = 0.5f / s;
        q.v[0] = (m.e[2][1] - m.e[1][2]) * s;
        q.v[1] = (m.e[0][2] - m.e[2][0]) * s;
        q.v[2] = (m.e[1][0] - m.e[0][1]) * s;
    } else {
        const int nxt[3] = {1, 2, 0};
        float qq[3];
        int i = 0;
        if (m.e[1][1] > m.e[0][0]) i = 1;
        if (m.e[2][2] > m.e[i][i]) i = 2;
        const int j = nxt[i], k = nxt[j];
        float s = std::sqrt((float)((double)(m.e[i][i] - (m.e[j][j] + m.e[k][k])) + 1.0));
        qq[i] = s * 0.5f;
        if (s != 0.f) s = 0.5f / s;
        q.w = (m.e[k][j] - m.e[j][k]) * s;
        qq[j] = (m.e[j][i] + m.e[i][j]) * s;
        qq[k] = (m.e[k][i] + m.e[i][k]) * s;
        q.v[0] = qq[0];
        q.v[1] = qq[1];
        q.v[2] = qq[2];
    }
    return q;
}

// Slerp (quaternion.cpp:73-84)
static Quat slerp(float t, const Quat& q1, const Quat& q2) {
    const float cos_theta = qdot(q1, q2);
    if (cos_theta > .9995f) return qnormalize(qadd(qmul(q1, 1.f - t), qmul(q2, t)));
    const float c = cos_theta < -1.f ? -1.f : cos_theta > 1.f ? 1.f : cos_theta;
    const float theta = std::acos(c);
    const float thetap = theta * t;
    const Quat qperp = qnormalize(qsub(q2, qmul(q1, cos_theta)));
    return qadd(qmul(q1, std::cos(thetap)), qmul(qperp, std::sin(thetap)));
}

// AnimatedTransform::Decompose (motionBlur.cpp:48-87): translation, rotation by
// polar decomposition (averaging with the inverse transpose until the change is
// at most 1e-4 or 100 steps), scale = inverse(R) * M
static void decompose(const Mat4& m, float T[3], Quat& Rq, Mat4& S) {
    T[0] = m.e[0][3];
    T[1] = m.e[1][3];
    T[2] = m.e[2][3];
    Mat4 M = m;
    for (int i = 0; i < 3; i++) M.e[i][3] = M.e[3][i] = 0.f;
    M.e[3][3] = 1.f;
    float norm;
    int count = 0;
    Mat4 R = M;
    do {
        Mat4 Rnext;
        const Mat4 Rit = inverse(transpose(R));
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) Rnext.e[i][j] = 0.5f * (R.e[i][j] + Rit.e[i][j]);
        norm = 0.f;
        for (int i = 0; i < 3; i++) {
            const float n = std::fabs(R.e[i][0] - Rnext.e[i][0]) + std::fabs(R.e[i][1] - Rnext.e[i][1]) +
                            std::fabs(R.e[i][2] - Rnext.e[i][2]);
            norm = std::max(norm, n);
        }
        R = Rnext;
    } while (++count < 100 && norm > .0001f);
    Rq = matrix_to_quat(R);
    S = mul(inverse(R), M);
}

Motion make_motion(float t0, float t1, const Mat4& a, const Mat4& b) {
    Motion m;
    m.t0 = t0;
    m.t1 = t1;
    m.start = a;
    m.end = b;
    decompose(a, m.T[0], m.R[0], m.S[0]);
    decompose(b, m.T[1], m.R[1], m.S[1]);
    bool equal = true;  // operator== over the 16 elements
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) equal = equal && a.e[i][j] == b.e[i][j];
    m.animated = !equal;
    return m;
}

// AnimatedTransform::Interpolate (motionBlur.cpp:6-45)
Mat4 interpolate(const Motion& m, float time) {
    if (!m.animated || time <= m.t0) return m.start;
    if (time >= m.t1) return m.end;
    const float dt = (time - m.t0) / (m.t1 - m.t0);
    float trans[3];
    for (int k = 0; k < 3; k++) trans[k] = (1.f - dt) * m.T[0][k] + dt * m.T[1][k];
    const Quat rotate = slerp(dt, m.R[0], m.R[1]);
    Mat4 scale = Mat4::identity();
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) scale.e[i][j] = (1.f - dt) * m.S[0].e[i][j] + dt * m.S[1].e[i][j];  // Lerp
    Mat4 I = Mat4::identity();
    for (int k = 0; k < 3; k++) I.e[k][3] = trans[k];  // set_translate
    return mul(mul(I, quat_to_matrix(rotate)), scale);
}

// ---- a small XML reader (the subset rapidxml's default parse gives the loader)
// Element names, attributes, and an element's value: the text of its first
// non-blank data run, untrimmed, with the five predefined entities and
// numeric character references decoded.  Comments, declarations, processing
// instructions and DOCTYPE are skipped.
struct XmlNode {
    std::string name, value;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XmlNode>> kids;

    const XmlNode* first(const char* n) const {
        for (const auto& k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
    const char* attr(const char* n) const {
        for (const auto& a : attrs)
            if (a.first == n) return a.second.c_str();
        return nullptr;
    }
    // value of child n, or nullptr
    const char* val(const char* n) const {
        const XmlNode* k = first(n);
        return k ? k->value.c_str() : nullptr;
    }
};

class XmlParser {
    const std::string& t;
    size_t p = 0;
    std::string where;

    [[noreturn]] void error(const std::string& msg) const {
        throw std::runtime_error(where + ": XML: " + msg + " at byte " + std::to_string(p));
    }
    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
    void skip_ws() {
        while (p < t.size() && ws(t[p])) p++;
    }
    bool starts(const char* s) const { return t.compare(p, strlen(s), s) == 0; }
    void skip_past(const char* s) {
        const size_t q = t.find(s, p);
        if (q == std::string::npos) error(std::string("missing '") + s + "'");
        p = q + strlen(s);
    }
    std::string decode(const std::string& s) const {
        std::string o;
        for (size_t i = 0; i < s.size(); i++) {
            if (s[i] != '&') {
                o += s[i];
                continue;
            }
            const size_t semi = s.find(';', i);
            if (semi == std::string::npos) {
                o += s[i];
                continue;
            }
            const std::string ent = s.substr(i + 1, semi - i - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "apos") o += '\'';
            else if (ent == "quot") o += '"';
            else if (!ent.empty() && ent[0] == '#') {
                const long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? strtol(ent.c_str() + 2, nullptr, 16)
                                                                                   : strtol(ent.c_str() + 1, nullptr, 10);
                if (c > 0 && c < 128) o += (char)c;  // ASCII only: the loader's values are ASCII
                else error("character reference outside ASCII");
            } else {
                o += s.substr(i, semi - i + 1);  // unknown entity kept verbatim
            }
            i = semi;
        }
        return o;
    }
    std::string name() {
        const size_t b = p;
        while (p < t.size() && !ws(t[p]) && t[p] != '>' && t[p] != '/' && t[p] != '=') p++;
        if (p == b) error("empty name");
        return t.substr(b, p - b);
    }
    // skips markup that is not an element; true if something was skipped
    bool skip_misc() {
        if (starts("<!--")) {
            skip_past("-->");
            return true;
        }
        if (starts("<?")) {
            skip_past("?>");
            return true;
        }
        if (starts("<![CDATA[")) error("CDATA sections are not supported");
        if (starts("<!")) {
            skip_past(">");
            return true;
        }
        return false;
    }
    std::unique_ptr<XmlNode> element() {
        // at '<'
        p++;
        auto n = std::make_unique<XmlNode>();
        n->name = name();
        for (;;) {
            skip_ws();
            if (p >= t.size()) error("unterminated start tag");
            if (t[p] == '/') {
                if (!starts("/>")) error("bad '/'");
                p += 2;
                return n;
            }
            if (t[p] == '>') {
                p++;
                break;
            }
            std::string an = name();
            skip_ws();
            if (p >= t.size() || t[p] != '=') error("attribute without value");
            p++;
            skip_ws();
            if (p >= t.size() || (t[p] != '"' && t[p] != '\'')) error("unquoted attribute");
            const char q = t[p++];
            const size_t e = t.find(q, p);
            if (e == std::string::npos) error("unterminated attribute");
            n->attrs.emplace_back(an, decode(t.substr(p, e - p)));
            p = e + 1;
        }
        for (;;) {  // contents
            const size_t start = p;
            skip_ws();
            if (p >= t.size()) error("unterminated element <" + n->name + ">");
            if (t[p] != '<') {  // a data run, from start (leading blanks kept) to the next '<'
                const size_t e = t.find('<', p);
                if (e == std::string::npos) error("unterminated element <" + n->name + ">");
                if (n->value.empty()) n->value = decode(t.substr(start, e - start));
                p = e;
                continue;
            }
            if (starts("</")) {
                p += 2;
                const std::string cn = name();
                if (cn != n->name) error("</" + cn + "> closes <" + n->name + ">");
                skip_ws();
                if (p >= t.size() || t[p] != '>') error("bad end tag");
                p++;
                return n;
            }
            if (skip_misc()) continue;
            n->kids.push_back(element());
        }
    }

public:
    XmlParser(const std::string& text, const std::string& file) : t(text), where(file) {}
    std::vector<std::unique_ptr<XmlNode>> document() {
        std::vector<std::unique_ptr<XmlNode>> top;
        for (;;) {
            skip_ws();
            if (p >= t.size()) return top;
            if (t[p] != '<') error("text outside an element");
            if (skip_misc()) continue;
            top.push_back(element());
        }
    }
};

// ---- value readers (Config::readVec / readMatrix, config.cpp:631-647) ------
static bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

static float atof_f(const char* s) { return (float)atof(s); }  // atof, then the float the setters take

static void read_vec(const char* s, float* v) {
    v[0] = v[1] = v[2] = 0.f;
    if (s) (void)sscanf(s, "(%f, %f, %f)", v, v + 1, v + 2);
}

// "[a, b, c, d][...][...][...]": rows of the matrix (set_value fills columns,
// then the reference transposes)
static Mat4 read_matrix(const char* s) {
    float v[16] = {0};
    (void)sscanf(s, "[%f, %f, %f, %f][%f, %f, %f, %f][%f, %f, %f, %f][%f, %f, %f, %f]", v, v + 1, v + 2, v + 3, v + 4,
                 v + 5, v + 6, v + 7, v + 8, v + 9, v + 10, v + 11, v + 12, v + 13, v + 14, v + 15);
    Mat4 m;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m.e[i][j] = v[4 * i + j];
    return m;
}

// ---- shapes (Mesh::loadShape, mesh.cpp:5-130) ------------------------------
static void bounding_box(const Shape& s, float* lo, float* hi) {  // Mesh::getBoundingBox (mesh.cpp:132-150)
    if (s.v.empty()) return;
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = s.v[a];
    for (size_t i = 1; i < s.v.size() / 3; i++)
        for (int a = 0; a < 3; a++) {
            const float x = s.v[3 * i + a];
            lo[a] = x < lo[a] ? x : lo[a];
            hi[a] = x > hi[a] ? x : hi[a];
        }
}

// Mesh::unitize (mesh.cpp:152-165): centres and scales the vertices in place
// and returns its matrix, whose translation the reference's column-major
// constructor puts in the bottom row (so applying it only scales).
static Mat4 unitize(Shape& s) {
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bounding_box(s, lo, hi);
    float c[3], d[3];
    for (int a = 0; a < 3; a++) {
        c[a] = (lo[a] + hi[a]) / 2.f;
        d[a] = hi[a] - lo[a];
    }
    float len = d[0] < d[1] ? d[1] : d[0];  // std::max
    len = len < d[2] ? d[2] : len;
    const float half = len / 2.f;
    for (size_t i = 0; i < s.v.size() / 3; i++)
        for (int a = 0; a < 3; a++) {
            s.v[3 * i + a] -= c[a];
            if (half != 0.f) s.v[3 * i + a] /= half;  // nvVector.h:375-379: a zero divisor is a no-op
        }
    Mat4 m = Mat4::identity();
    for (int a = 0; a < 3; a++) {
        m.e[a][a] = 2.f / len;
        m.e[3][a] = -2.f / len * c[a];
    }
    return m;
}

// Mesh::applyTransform (mesh.cpp:167-174)
static void apply_transform(Shape& s, const Mat4& t) {
    const Mat4 nm = transpose(inverse(t));
    float o[3];
    for (size_t i = 0; i < s.v.size() / 3; i++) {
        t.apply(&s.v[3 * i], 1.f, o);
        for (int a = 0; a < 3; a++) s.v[3 * i + a] = o[a];
    }
    for (size_t i = 0; i < s.vn.size() / 3; i++) {
        nm.apply(&s.vn[3 * i], 0.f, o);
        for (int a = 0; a < 3; a++) s.vn[3 * i + a] = o[a];
    }
}

static const int kLineBuffer = 1024;  // BUFFERSIZE (macros.h:7): fgets splits longer lines

void load_shape(const std::string& file, bool normalize, Shape& s, std::vector<Shape>* split) {
    s.v.clear();
    s.f.clear();
    if (file.size() < 4) throw std::runtime_error(file + ": not a .obj or .shp file");
    const std::string suffix = file.substr(file.size() - 4);
    if (suffix == ".shp") {
        FILE* fp = fopen(file.c_str(), "rb");
        if (!fp) throw std::runtime_error(file + ": cannot open");
        int32_t n = 0;
        bool ok = fread(&n, 4, 1, fp) == 1 && n >= 0;
        // the reference sizes the vertex list to n vec3s but reads n floats into it
        // (mesh.cpp:20-25): the first n/3 vertices hold the data, the rest are 0
        if (ok) {
            s.v.assign(3 * (size_t)n, 0.f);
            ok = fread(s.v.data(), 4, (size_t)n, fp) == (size_t)n;
        }
        if (ok) ok = fread(&n, 4, 1, fp) == 1 && n >= 0;
        if (ok) {
            s.f.assign(3 * (size_t)n, 0u);
            ok = fread(s.f.data(), 4, (size_t)n, fp) == (size_t)n;
        }
        fclose(fp);
        if (!ok) throw std::runtime_error(file + ": truncated .shp file");
    } else if (suffix == ".obj") {
        FILE* fp = fopen(file.c_str(), "r");
        if (!fp) throw std::runtime_error(file + ": cannot open");
        char line[kLineBuffer], attrib[kLineBuffer], parm[3][kLineBuffer];
        size_t cur_v = 0, cur_n = 0;
        Shape* shape = nullptr;
        while (fgets(line, kLineBuffer, fp)) {
            if (line[0] == '#') continue;
            attrib[0] = 0;
            const int num = sscanf(line, "%s %s %s %s", attrib, parm[0], parm[1], parm[2]);
            if (strcmp("g", attrib) == 0 && split) {
                split->emplace_back();
                shape = &split->back();
                shape->name = num >= 2 ? parm[0] : "";
                cur_v = s.v.size() / 3;
                cur_n = s.vn.size() / 3;
            }
            if (num != 4) continue;
            if (strcmp("v", attrib) == 0 || strcmp("vn", attrib) == 0) {
                float x[3] = {0, 0, 0};
                for (int k = 0; k < 3; k++) (void)sscanf(parm[k], "%f", &x[k]);
                std::vector<float>& dst = attrib[1] ? s.vn : s.v;
                dst.insert(dst.end(), x, x + 3);
                if (split && shape) {
                    std::vector<float>& sd = attrib[1] ? shape->vn : shape->v;
                    sd.insert(sd.end(), x, x + 3);
                }
            }
            if (strcmp("f", attrib) == 0) {
                unsigned tri[3] = {0, 0, 0}, tex[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
                int ret = 0;
                for (int k = 0; k < 3; k++) ret = sscanf(parm[k], "%u/%u/%u", &tri[k], &tex[k], &nrm[k]);
                if (ret == 1)
                    for (int k = 0; k < 3; k++) ret = sscanf(parm[k], "%u//%u", &tri[k], &nrm[k]);
                if (ret == 2 && strstr(parm[2], "//") == nullptr)
                    throw std::runtime_error(file + ": faces with texture but no normal indices (v/vt) leave the "
                                                    "reference's normal indices undefined");
                for (int k = 0; k < 3; k++) {
                    tri[k] -= 1u;
                    nrm[k] -= 1u;
                }
                s.f.insert(s.f.end(), tri, tri + 3);
                if (ret >= 2) s.fn.insert(s.fn.end(), nrm, nrm + 3);
                if (split && shape) {
                    for (int k = 0; k < 3; k++) shape->f.push_back(tri[k] - (unsigned)cur_v);
                    if (ret >= 2)
                        for (int k = 0; k < 3; k++) shape->fn.push_back(nrm[k] - (unsigned)cur_n);
                }
            }
        }
        fclose(fp);
    } else {
        throw std::runtime_error(file + ": unknown shape suffix '" + suffix + "' (.obj and .shp are read)");
    }
    if (normalize) {
        const Mat4 um = unitize(s);
        if (split)
            for (Shape& sh : *split) {
                apply_transform(sh, um);
                sh.transform = s.transform;
            }
    }
}

// ---- scene configuration (Config::load, config.cpp:432-575) -----------------
struct ConfigLoader {
    std::string root, current;
    std::unordered_map<std::string, std::vector<std::unique_ptr<XmlNode>>> docs;
    Scene& out;

    explicit ConfigLoader(Scene& o) : out(o) {}

    std::string resolve(const std::string& p) const { return file_exists(p) ? p : root + "/" + p; }

    const std::vector<std::unique_ptr<XmlNode>>& doc(const std::string& file) {
        const std::string full = resolve(file);
        auto it = docs.find(full);
        if (it != docs.end()) return it->second;
        FILE* fp = fopen(full.c_str(), "rb");
        if (!fp) throw std::runtime_error(full + ": cannot open");
        std::string text;
        char buf[65536];
        size_t n;
        while ((n = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
        fclose(fp);
        return docs[full] = XmlParser(text, full).document();
    }

    // Config::findNode (config.cpp:648-671): the first top-level node with the
    // tag whose Name attribute matches (or that has none, or any when name is "")
    const XmlNode* find(const std::string& file, const char* tag, const std::string& name) {
        if (file.empty()) return nullptr;
        for (const auto& n : doc(file)) {
            if (n->name != tag) continue;
            const char* a = n->attr("Name");
            if (!a || name.empty() || name == a) return n.get();
        }
        return nullptr;
    }
    static const XmlNode* find_child(const XmlNode* root, const char* tag, const std::string& name) {
        for (const auto& n : root->kids) {
            if (n->name != tag) continue;
            const char* a = n->attr("Name");
            if (!a || name.empty() || name == a) return n.get();
        }
        return nullptr;
    }
    std::pair<std::string, std::string> path_name(const XmlNode* n) const {  // getPathAndName (config.cpp:672-677)
        const char* p = n->val("path");
        const char* m = n->val("name");
        return {p ? p : current, m ? m : ""};
    }

    static const char* need(const XmlNode* n, const char* child, const std::string& ctx) {
        const char* v = n->val(child);
        if (!v) throw std::runtime_error(ctx + ": <" + child + "> is missing");
        return v;
    }

    // AbstractObject from an <Object> and its <Material> (Config::generateSceneObject)
    void object(const std::string& shape_path, const XmlNode* obj, const XmlNode* mat, const Shape* shape,
                const std::string& ctx) {
        if (!mat) throw std::runtime_error(ctx + ": material not found");
        const std::string type = need(mat, "type", ctx + " material");
        SceneObject o;
        o.mat = wr_material{};
        if (type == "Diffuse") {
            o.mat.type = WR_DIFFUSE;
            read_vec(need(mat, "color", ctx), o.mat.color);
        } else if (type == "Glossy") {
            o.mat.type = WR_GLOSSY;
            read_vec(need(mat, "color", ctx), o.mat.color);
            o.mat.power = atof_f(need(mat, "power", ctx));
        } else if (type == "Emissive") {
            o.mat.type = WR_EMISSIVE;
            read_vec(need(mat, "color", ctx), o.mat.color);
        } else if (type == "Refractive") {
            o.mat.type = WR_REFRACTIVE;
            o.mat.ior = 1.5f;  // refractiveObj.h:10
            if (const char* v = mat->val("IOR")) o.mat.ior = atof_f(v);
            if (const char* v = mat->val("decayColor")) {
                float dc[3];
                read_vec(v, dc);
                if (dc[0] != 1.f || dc[1] != 1.f || dc[2] != 1.f)
                    throw std::runtime_error(ctx + ": refractive decayColor other than (1, 1, 1) is not supported");
            }
        } else {
            throw std::runtime_error(ctx + ": object class '" + type + "' is not on the GPU path");
        }
        if (const XmlNode* mo = obj->first("motion")) {  // config.cpp:28-45
            o.moving = true;
            o.t0 = atof_f(need(mo, "startTime", ctx));
            o.t1 = atof_f(need(mo, "endTime", ctx));
            o.m0 = read_matrix(need(mo, "startTransform", ctx));
            o.m1 = read_matrix(need(mo, "endTransform", ctx));
        }
        Mat4 transform = Mat4::identity();
        if (const char* t = obj->val("transform")) transform = read_matrix(t);
        Shape sh;
        if (shape) {
            sh = *shape;  // copyMesh: the group's shape, and its transform replaces the object's
            transform = shape->transform;
        } else {
            load_shape(shape_path, true, sh, nullptr);
        }
        const size_t nv = sh.v.size() / 3, nt = sh.f.size() / 3;
        for (uint32_t x : sh.f)
            if (x >= nv) throw std::runtime_error(ctx + ": face vertex index out of range");
        // a moving object's own transform is replaced before every pass
        // (AbstractObject::updateTransform): keep its mesh in its own space
        if (o.moving) transform = Mat4::identity();
        // world vertices: getWorldVertexPosition (mesh.h:106)
        o.verts.resize(3 * nv);
        for (size_t i = 0; i < nv; i++) transform.apply(&sh.v[3 * i], 1.f, &o.verts[3 * i]);
        o.tris = sh.f;
        // world vertex normals per triangle corner, as Mesh::getWorldNormal forms them (mesh.h:47-77)
        if (!sh.vn.empty()) {
            if (sh.fn.size() != sh.f.size())
                throw std::runtime_error(ctx + ": vertex normals without a normal index for every face");
            const Mat4 nm = transpose(inverse(transform));
            o.corner_normals.resize(9 * nt);
            for (size_t k = 0; k < 3 * nt; k++) {
                if (sh.fn[k] >= sh.vn.size() / 3) throw std::runtime_error(ctx + ": normal index out of range");
                nm.apply(&sh.vn[3 * (size_t)sh.fn[k]], 0.f, &o.corner_normals[3 * k]);
            }
        }
        o.name = obj->attr("Name") ? obj->attr("Name") : "";
        out.objects.push_back(std::move(o));
    }

    void load(const std::string& path) {
        const size_t p1 = path.rfind('/'), p2 = path.rfind('\\');
        const size_t pos = std::min(p1, p2);
        root = pos == std::string::npos ? std::string(".") : path.substr(0, pos);
        const XmlNode* cfg = find(path, "Config", "");
        if (!cfg) throw std::runtime_error(path + ": no <Config>");
        const XmlNode* scfg = cfg->first("Scene");
        if (!scfg) throw std::runtime_error(path + ": <Config> has no <Scene>");
        auto pn = path_name(scfg);
        const XmlNode* scene = find(pn.first, "Scene", pn.second);
        if (!scene) scene = find(path, "Scene", pn.second);
        if (!scene) throw std::runtime_error(path + ": scene '" + pn.second + "' not found");
        out.name = pn.second;
        current = pn.first;
        for (const auto& g : scene->kids) {  // GroupObject: one .obj split into named shapes
            if (g->name != "GroupObject") continue;
            const std::string file = resolve(need(g.get(), "filePath", "GroupObject"));
            Shape whole;
            if (const char* t = g->val("transform")) whole.transform = read_matrix(t);
            std::vector<Shape> shapes;
            load_shape(file, true, whole, &shapes);
            std::unordered_map<std::string, const Shape*> by_name;
            for (const Shape& s : shapes) by_name[s.name] = &s;
            for (const auto& o : g->kids) {
                if (o->name != "Object") continue;
                const XmlNode* mref = o->first("Material");
                if (!mref) throw std::runtime_error(file + ": <Object> without <Material>");
                auto mpn = path_name(mref);
                const char* nm = o->attr("Name");
                auto it = by_name.find(nm ? nm : "");
                if (it == by_name.end()) throw std::runtime_error(file + ": no group named '" + (nm ? nm : "") + "'");
                object(file, o.get(), find(mpn.first, "Material", mpn.second), it->second,
                       file + " object " + (nm ? nm : ""));
            }
        }
        for (const auto& o : scene->kids) {
            if (o->name != "Object") continue;
            const XmlNode* mref = o->first("Material");
            if (!mref) throw std::runtime_error(path + ": <Object> without <Material>");
            auto mpn = path_name(mref);
            const std::string file = resolve(need(o.get(), "filePath", "Object"));
            object(file, o.get(), find(mpn.first, "Material", mpn.second), nullptr, file);
        }
        const char* camname = scfg->val("camera");
        const XmlNode* cam = find_child(scene, "Camera", camname ? camname : "");
        if (!cam) throw std::runtime_error(path + ": camera not found");
        read_vec(need(cam, "position", "Camera"), out.cam.pos);
        read_vec(need(cam, "focus", "Camera"), out.cam.focus);
        read_vec(need(cam, "up", "Camera"), out.cam.up);
        out.cam.width = (uint32_t)atof_f(need(cam, "width", "Camera"));
        out.cam.height = (uint32_t)atof_f(need(cam, "height", "Camera"));
        out.cam.sight_dist = atof_f(need(cam, "sightDist", "Camera"));
        if (cam->first("lensRadius") && cam->first("focalDist"))
            throw std::runtime_error(path + ": depth of field (lensRadius, focalDist) is not supported");
        out.renderer = "";
        out.spp = 0;
        out.max_len = 0;
        if (const char* r = cfg->val("renderer")) {
            out.renderer = r;
            out.spp = INT_MAX;  // config.cpp:521-527
            if (const char* it = cfg->val("iter")) out.spp = (int)atof(it);
            out.max_len = 20;
            if (const char* rd = cfg->val("radius")) out.radius = (float)atof(rd);  // setRadius(atof(...))
        }
    }
};

void load_config(const std::string& path, Scene& out) {
    ConfigLoader L(out);
    L.load(path);
}

// Film::savePFM after Film::setBuffer: header "PF\n<W> <H>\n-1\n", then the
// rows bottom to top as native (little-endian) floats.
void write_pfm(const std::string& path, const float* rgb, uint32_t w, uint32_t h) {
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error(path + ": cannot create");
    const std::string hdr = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1\n";
    bool ok = fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
    for (uint32_t y = 0; ok && y < h; y++) {
        const float* row = rgb + 3 * (size_t)(h - 1 - y) * w;
        ok = fwrite(row, sizeof(float), 3 * (size_t)w, fp) == 3 * (size_t)w;
    }
    if (fclose(fp) != 0) ok = false;
    if (!ok) throw std::runtime_error(path + ": write failed");
}

}  // namespace io
}  // namespace wr

// ---- C ABI -------------------------------------------------------------------
namespace wr {
int fail(wr_ctx* ctx, int code, const std::string& msg);  // capi.cpp: records the message for wr_last_error
}

struct wr_config {
    wr::io::Scene scene;
};
struct wr_shape {
    wr::io::Shape shape;
};

extern "C" {

int wr_config_load(const char* path, wr_config** out) {
    if (!path || !out) return wr::fail(nullptr, WR_EINVAL, "wr_config_load: bad argument");
    *out = nullptr;
    try {
        std::unique_ptr<wr_config> c(new wr_config);
        wr::io::load_config(path, c->scene);
        *out = c.release();
        return WR_OK;
    } catch (const std::exception& e) {
        return wr::fail(nullptr, WR_EINVAL, std::string("wr_config_load: ") + e.what());
    }
}

void wr_config_free(wr_config* cfg) { delete cfg; }

int wr_config_get_info(const wr_config* cfg, wr_config_info* info) {
    if (!cfg || !info) return wr::fail(nullptr, WR_EINVAL, "wr_config_get_info: bad argument");
    const wr::io::Scene& s = cfg->scene;
    *info = wr_config_info{};
    info->n_objects = (uint32_t)s.objects.size();
    info->spp = s.spp;
    info->max_len = s.max_len;
    info->radius = s.radius;
    info->camera = s.cam;
    snprintf(info->renderer, sizeof info->renderer, "%s", s.renderer.c_str());
    snprintf(info->scene_name, sizeof info->scene_name, "%s", s.name.c_str());
    return WR_OK;
}

int wr_config_get_object(const wr_config* cfg, uint32_t i, uint32_t* nv, uint32_t* nt, uint32_t* has_normals,
                         float* verts, uint32_t* tris, float* corner_normals, wr_material* mat) {
    if (!cfg || !nv || !nt || !has_normals) return wr::fail(nullptr, WR_EINVAL, "wr_config_get_object: bad argument");
    if (i >= cfg->scene.objects.size()) return wr::fail(nullptr, WR_EINVAL, "wr_config_get_object: no such object");
    const wr::io::SceneObject& o = cfg->scene.objects[i];
    *nv = (uint32_t)(o.verts.size() / 3);
    *nt = (uint32_t)(o.tris.size() / 3);
    *has_normals = o.corner_normals.empty() ? 0u : 1u;
    if (verts) memcpy(verts, o.verts.data(), o.verts.size() * sizeof(float));
    if (tris) memcpy(tris, o.tris.data(), o.tris.size() * sizeof(uint32_t));
    if (corner_normals && !o.corner_normals.empty())
        memcpy(corner_normals, o.corner_normals.data(), o.corner_normals.size() * sizeof(float));
    if (mat) *mat = o.mat;
    return WR_OK;
}

int wr_config_add_to_scene(const wr_config* cfg, wr_scene* scene) {
    if (!cfg || !scene) return wr::fail(nullptr, WR_EINVAL, "wr_config_add_to_scene: bad argument");
    for (const wr::io::SceneObject& o : cfg->scene.objects) {
        uint32_t id = 0;
        int rc = wr_scene_add_mesh_corners(scene, o.verts.data(), (uint32_t)(o.verts.size() / 3), o.tris.data(),
                                           (uint32_t)(o.tris.size() / 3),
                                           o.corner_normals.empty() ? nullptr : o.corner_normals.data(), &o.mat, &id);
        if (rc != WR_OK) return rc;
        if (o.moving) {
            float a[16], b[16];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    a[4 * i + j] = o.m0.e[i][j];
                    b[4 * i + j] = o.m1.e[i][j];
                }
            rc = wr_scene_set_motion(scene, id, o.t0, o.t1, a, b);
            if (rc != WR_OK) return rc;
        }
    }
    return WR_OK;
}

int wr_config_get_motion(const wr_config* cfg, uint32_t i, int* moving, float* start_time, float* end_time,
                         float* start, float* end) {
    if (!cfg || !moving || i >= cfg->scene.objects.size())
        return wr::fail(nullptr, WR_EINVAL, "wr_config_get_motion: bad argument");
    const wr::io::SceneObject& o = cfg->scene.objects[i];
    *moving = o.moving ? 1 : 0;
    if (start_time) *start_time = o.t0;
    if (end_time) *end_time = o.t1;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            if (start) start[4 * r + c] = o.m0.e[r][c];
            if (end) end[4 * r + c] = o.m1.e[r][c];
        }
    return WR_OK;
}

int wr_shape_load(const char* path, int normalize, wr_shape** out) {
    if (!path || !out) return wr::fail(nullptr, WR_EINVAL, "wr_shape_load: bad argument");
    *out = nullptr;
    try {
        std::unique_ptr<wr_shape> s(new wr_shape);
        wr::io::load_shape(path, normalize != 0, s->shape, nullptr);
        *out = s.release();
        return WR_OK;
    } catch (const std::exception& e) {
        return wr::fail(nullptr, WR_EINVAL, std::string("wr_shape_load: ") + e.what());
    }
}

void wr_shape_free(wr_shape* s) { delete s; }

int wr_shape_get(const wr_shape* s, uint32_t* nv, uint32_t* nt, uint32_t* nn, uint32_t* nfn, float* verts,
                 uint32_t* tris, float* normals, uint32_t* normal_tris) {
    if (!s || !nv || !nt || !nn || !nfn) return wr::fail(nullptr, WR_EINVAL, "wr_shape_get: bad argument");
    const wr::io::Shape& sh = s->shape;
    *nv = (uint32_t)(sh.v.size() / 3);
    *nt = (uint32_t)(sh.f.size() / 3);
    *nn = (uint32_t)(sh.vn.size() / 3);
    *nfn = (uint32_t)(sh.fn.size() / 3);
    if (verts) memcpy(verts, sh.v.data(), sh.v.size() * sizeof(float));
    if (tris) memcpy(tris, sh.f.data(), sh.f.size() * sizeof(uint32_t));
    if (normals) memcpy(normals, sh.vn.data(), sh.vn.size() * sizeof(float));
    if (normal_tris) memcpy(normal_tris, sh.fn.data(), sh.fn.size() * sizeof(uint32_t));
    return WR_OK;
}

int wr_motion_interpolate(float start_time, float end_time, const float* start, const float* end, float time,
                          float* out) {
    if (!start || !end || !out) return wr::fail(nullptr, WR_EINVAL, "wr_motion_interpolate: bad argument");
    wr::io::Mat4 a, b;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a.e[i][j] = start[4 * i + j];
            b.e[i][j] = end[4 * i + j];
        }
    const wr::io::Mat4 t = wr::io::interpolate(wr::io::make_motion(start_time, end_time, a, b), time);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = t.e[i][j];
    return WR_OK;
}

int wr_pfm_write(const char* path, const float* rgb, uint32_t width, uint32_t height) {
    if (!path || (!rgb && width && height)) return wr::fail(nullptr, WR_EINVAL, "wr_pfm_write: bad argument");
    try {
        wr::io::write_pfm(path, rgb, width, height);
        return WR_OK;
    } catch (const std::exception& e) {
        return wr::fail(nullptr, WR_EINVAL, std::string("wr_pfm_write: ") + e.what());
    }
}

}  // extern "C"

// wr_sceneio.h -- host-side scene input/output (sceneio.cpp), internal types.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/walkray.h"

namespace wr {
namespace io {

// 4x4 float matrix indexed e[row][col] (nvMatrix.h's element(row, col)).
struct Mat4 {
    float e[4][4];
    static Mat4 identity();
    // (M * (v, w)).xyz with each row summed left to right
    void apply(const float* v, float w, float* out) const;
};
Mat4 transpose(const Mat4& m);
Mat4 inverse(const Mat4& m);
Mat4 mul(const Mat4& a, const Mat4& b);

// AnimatedTransform (motionBlur.h:6-34): the two transforms, their times and
// their decompositions (translation, rotation quaternion, scale matrix).
struct Quat {
    float v[3];
    float w;
};
struct Motion {
    float t0 = 0.f, t1 = 1.f;
    Mat4 start = Mat4::identity(), end = Mat4::identity();
    bool animated = false;  // actuallyAnimated
    float T[2][3] = {};
    Quat R[2] = {};
    Mat4 S[2] = {Mat4::identity(), Mat4::identity()};
};
Motion make_motion(float t0, float t1, const Mat4& start, const Mat4& end);
// AnimatedTransform::Interpolate (motionBlur.cpp:6-45)
Mat4 interpolate(const Motion& m, float time);

// Mesh's lists (mesh.h:13-19): positions, vertex normals, and per-face
// position / normal indices (3 per face), plus the mesh transform.
struct Shape {
    std::string name;
    std::vector<float> v, vn;
    std::vector<uint32_t> f, fn;
    Mat4 transform = Mat4::identity();
};

// Mesh::loadShape (mesh.cpp:5-130).  split != nullptr: also the "g" groups
// of an .obj as separate shapes.
void load_shape(const std::string& file, bool normalize, Shape& s, std::vector<Shape>* split);

// One loaded object in world space: vertices, triangles, per-corner world
// normals (9 floats per triangle, empty for flat shading), material.
struct SceneObject {
    std::string name;
    std::vector<float> verts;
    std::vector<uint32_t> tris;
    std::vector<float> corner_normals;
    wr_material mat;
    // <motion> (config.cpp:28-45): verts / corner_normals are then the mesh's
    // own space, placed per pass by the AnimatedTransform of these
    bool moving = false;
    float t0 = 0.f, t1 = 0.f;
    Mat4 m0 = Mat4::identity(), m1 = Mat4::identity();
};

struct Scene {
    std::string name;
    std::vector<SceneObject> objects;
    wr_camera cam{};
    std::string renderer;  // empty when the configuration names none
    int spp = 0;
    int max_len = 0;
    float radius = 0.f;  // <radius> of VCMTracer / PhotonMap (config.cpp:560,571), 0 if absent
};

// Config::load (config.cpp:432-575).  Throws std::runtime_error with a message.
void load_config(const std::string& path, Scene& out);

void write_pfm(const std::string& path, const float* rgb, uint32_t w, uint32_t h);

}  // namespace io
}  // namespace wr

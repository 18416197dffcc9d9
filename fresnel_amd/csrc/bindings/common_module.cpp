// _common: plain data types shared by the Python API and the `_gpu` module.
// Mirrors the reference's `_common` module (fresnel/common/module-common.cc:37-115);
// the C++ types are the C-ABI structs of include/fresnel_hip.h.
#include <pybind11/pybind11.h>
#include <cmath>
#include <sstream>

#include "fresnel_hip.h"

namespace py = pybind11;

enum class CameraModel
    {
    orthographic = 0,
    perspective = 1
    };

// Orthonormal camera basis u (right), v (up), w (back), in the float operation
// order of CameraBasis (fresnel/common/Camera.h:92-110).
struct CameraBasis
    {
    fh_vec3 u, v, w;
    explicit CameraBasis(const fh_user_camera& c)
        {
        auto dot = [](fh_vec3 a, fh_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
        auto cross = [](fh_vec3 a, fh_vec3 b)
        { return fh_vec3 {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
        auto scale = [](fh_vec3 a, float s) { return fh_vec3 {a.x * s, a.y * s, a.z * s}; };
        fh_vec3 d {c.look_at.x - c.position.x, c.look_at.y - c.position.y, c.look_at.z - c.position.z};
        fh_vec3 vv = c.up;
        d = scale(d, 1.0f / sqrtf(dot(d, d)));
        vv = scale(vv, 1.0f / sqrtf(dot(vv, vv)));
        fh_vec3 uu = cross(d, vv);
        uu = scale(uu, 1.0f / sqrtf(dot(uu, uu)));
        vv = cross(uu, d);
        vv = scale(vv, 1.0f / sqrtf(dot(vv, vv)));
        u = uu;
        v = vv;
        w = fh_vec3 {-d.x, -d.y, -d.z};
        }
    };

static fh_user_camera default_camera()
    {
    // UserCamera(), fresnel/common/Camera.h:57-67
    fh_user_camera c;
    c.position = {0, 0, 0};
    c.look_at = {0, 0, 1};
    c.up = {0, 1, 0};
    c.h = 1;
    c.f = 1;
    c.f_stop = 1;
    c.focus_distance = 1;
    c.model = 0;
    return c;
    }

PYBIND11_MODULE(_common, m)
    {
    m.def("gpu_built", []() { return true; });
    m.def("cpu_built", []() { return false; });

    py::class_<fh_rgb>(m, "RGBf")
        .def(py::init([](float r, float g, float b) { return fh_rgb {r, g, b}; }))
        .def_readwrite("r", &fh_rgb::r)
        .def_readwrite("g", &fh_rgb::g)
        .def_readwrite("b", &fh_rgb::b);

    py::class_<fh_material>(m, "Material")
        .def(py::init([]() { return fh_material {0, {0, 0, 0}, 0, 0, 0, 0, 0}; }))
        .def_readwrite("solid", &fh_material::solid)
        .def_readwrite("primitive_color_mix", &fh_material::primitive_color_mix)
        .def_readwrite("roughness", &fh_material::roughness)
        .def_readwrite("specular", &fh_material::specular)
        .def_readwrite("spec_trans", &fh_material::spec_trans)
        .def_readwrite("metal", &fh_material::metal)
        .def_readwrite("color", &fh_material::color)
        .def("__repr__",
             [](const fh_material& a)
             {
                 std::ostringstream s;
                 s << "<fresnel._common.Material:"
                   << " solid=" << a.solid << " color=(" << a.color.r << ", " << a.color.g << ", "
                   << a.color.b << ")"
                   << " primitive_color_mix=" << a.primitive_color_mix
                   << " roughness=" << a.roughness << " specular=" << a.specular
                   << " spec_trans=" << a.spec_trans << " metal=" << a.metal << ">";
                 return s.str();
             });

    py::class_<fh_vec3>(m, "vec3f")
        .def(py::init([](float x, float y, float z) { return fh_vec3 {x, y, z}; }))
        .def_readwrite("x", &fh_vec3::x)
        .def_readwrite("y", &fh_vec3::y)
        .def_readwrite("z", &fh_vec3::z)
        .def("__repr__",
             [](const fh_vec3& a)
             {
                 std::ostringstream s;
                 s << "<fresnel._common.vec3f (" << a.x << ", " << a.y << ", " << a.z << ")>";
                 return s.str();
             });

    py::enum_<CameraModel>(m, "CameraModel")
        .value("orthographic", CameraModel::orthographic)
        .value("perspective", CameraModel::perspective);

    py::class_<fh_user_camera>(m, "UserCamera")
        .def(py::init([]() { return default_camera(); }))
        .def_readwrite("position", &fh_user_camera::position)
        .def_readwrite("look_at", &fh_user_camera::look_at)
        .def_readwrite("up", &fh_user_camera::up)
        .def_readwrite("h", &fh_user_camera::h)
        .def_readwrite("f", &fh_user_camera::f)
        .def_readwrite("f_stop", &fh_user_camera::f_stop)
        .def_readwrite("focus_distance", &fh_user_camera::focus_distance)
        .def_property(
            "model",
            [](const fh_user_camera& c) { return static_cast<CameraModel>(c.model); },
            [](fh_user_camera& c, CameraModel v) { c.model = static_cast<int32_t>(v); });

    py::class_<CameraBasis>(m, "CameraBasis")
        .def(py::init<const fh_user_camera&>())
        .def_readwrite("u", &CameraBasis::u)
        .def_readwrite("v", &CameraBasis::v)
        .def_readwrite("w", &CameraBasis::w);

    py::class_<fh_lights>(m, "Lights")
        .def(py::init(
            []()
            {
                fh_lights l;
                for (int i = 0; i < 4; i++)
                    {
                    l.direction[i] = {0, 0, 0};
                    l.color[i] = {0, 0, 0};
                    l.theta[i] = 0;
                    }
                l.N = 0;
                return l;
            }))
        .def_readwrite("N", &fh_lights::N)
        .def("getDirection", [](fh_lights& l, unsigned i) { return l.direction[i & 3]; })
        .def("setDirection", [](fh_lights& l, unsigned i, const fh_vec3& v) { l.direction[i & 3] = v; })
        .def("getColor", [](fh_lights& l, unsigned i) { return l.color[i & 3]; })
        .def("setColor", [](fh_lights& l, unsigned i, const fh_rgb& v) { l.color[i & 3] = v; })
        .def("getTheta", [](fh_lights& l, unsigned i) { return l.theta[i & 3]; })
        .def("setTheta", [](fh_lights& l, unsigned i, float v) { l.theta[i & 3] = v; });
    }

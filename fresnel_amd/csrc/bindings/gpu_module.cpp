// _gpu: pybind11 binding of libfresnel_hip.so with the class set of the
// reference's `_gpu` module (fresnel/gpu/module-gpu.cc:22-43): Device, Scene,
// Geometry + GeometrySphere/Cylinder/Polygon/ConvexPolyhedron/Mesh, Tracer,
// TracerDirect, TracerPath and buffer arrays with map()/unmap().
//
// Every call goes through the C ABI of include/fresnel_hip.h; a non-zero
// status becomes RuntimeError with the library's message.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fresnel_hip.h"

namespace py = pybind11;

static void check(int status)
    {
    if (status != FH_OK)
        throw std::runtime_error(fh_last_error());
    }

struct Device
    {
    fh_device h = nullptr;
    Device(const std::string& /*code object directory*/, int limit)
        {
        check(fh_device_create(limit, &h));
        }
    ~Device()
        {
        fh_device_destroy(h);
        }
    std::string describe()
        {
        char buf[1024];
        check(fh_device_describe(h, buf, sizeof(buf)));
        return buf;
        }
    static std::string getAllGPUs()
        {
        char buf[4096];
        check(fh_device_get_all_gpus(buf, sizeof(buf)));
        return buf;
        }
    };

struct Scene
    {
    std::shared_ptr<Device> dev;
    fh_scene h = nullptr;
    fh_user_camera camera {};
    fh_lights lights {};
    fh_rgb bg {0, 0, 0};
    float bg_alpha = 0;
    explicit Scene(std::shared_ptr<Device> d) : dev(d)
        {
        check(fh_scene_create(d->h, &h));
        check(fh_scene_get_camera(h, &camera));
        check(fh_scene_get_lights(h, &lights));
        check(fh_scene_get_background(h, &bg, &bg_alpha));
        }
    ~Scene()
        {
        fh_scene_destroy(h);
        }
    // camera and lights are edited in place through references, as with the
    // reference's reference_internal getters; push them before each render
    void push()
        {
        check(fh_scene_set_camera(h, &camera));
        check(fh_scene_set_lights(h, &lights));
        check(fh_scene_set_background(h, bg, bg_alpha));
        }
    };

// Host view of a backend buffer.  For geometry buffers the host copy is
// authoritative; for tracer outputs map() copies the device image first.
struct __attribute__((visibility("hidden"))) Array
    {
    std::function<void()> on_map;
    std::function<py::buffer_info()> info;
    std::function<void(uint64_t, uint64_t)> on_rows;
    std::function<py::tuple()> layout; // (shape, numpy dtype string) without a copy
    std::shared_ptr<void> owner;
    void map()
        {
        if (on_map)
            on_map();
        }
    void unmap() { }
    // rows [first, first + count) of a geometry buffer were written
    void update_rows(uint64_t first, uint64_t count)
        {
        if (!on_rows)
            throw std::runtime_error("not a geometry buffer");
        on_rows(first, count);
        }
    };

struct Geometry
    {
    std::shared_ptr<Scene> scene;
    fh_geometry h = nullptr;
    virtual ~Geometry()
        {
        fh_geometry_destroy(h);
        }
    fh_material getMaterial()
        {
        fh_material m;
        check(fh_geometry_get_material(h, &m));
        return m;
        }
    void setMaterial(const fh_material& m)
        {
        check(fh_geometry_set_material(h, &m));
        }
    fh_material getOutlineMaterial()
        {
        fh_material m;
        check(fh_geometry_get_outline_material(h, &m));
        return m;
        }
    void setOutlineMaterial(const fh_material& m)
        {
        check(fh_geometry_set_outline_material(h, &m));
        }
    float getOutlineWidth()
        {
        float w;
        check(fh_geometry_get_outline_width(h, &w));
        return w;
        }
    void setOutlineWidth(float w)
        {
        check(fh_geometry_set_outline_width(h, w));
        }
    void enable()
        {
        check(fh_geometry_enable(h));
        }
    void disable()
        {
        check(fh_geometry_disable(h));
        }
    void remove()
        {
        check(fh_geometry_remove(h));
        }
    void update()
        {
        check(fh_geometry_update(h));
        }
    };

static std::shared_ptr<Array> geometry_buffer(std::shared_ptr<Geometry> g, int which)
    {
    float* data = nullptr;
    uint32_t ndim = 0;
    uint64_t shape[3];
    check(fh_geometry_buffer(g->h, which, &data, &ndim, shape));
    auto a = std::make_shared<Array>();
    a->owner = g;
    std::vector<py::ssize_t> shp(shape, shape + ndim);
    std::vector<py::ssize_t> strides(ndim);
    py::ssize_t st = sizeof(float);
    for (int k = (int)ndim - 1; k >= 0; k--)
        {
        strides[k] = st;
        st *= shp[k];
        }
    a->info = [data, shp, strides]()
    { return py::buffer_info(data, sizeof(float), py::format_descriptor<float>::format(), (py::ssize_t)shp.size(), shp, strides); };
    a->layout = [shp]() { return py::make_tuple(py::tuple(py::cast(shp)), "float32"); };
    Geometry* gp = g.get();
    a->on_rows = [gp, which](uint64_t first, uint64_t count) { check(fh_geometry_update_rows(gp->h, which, first, count)); };
    return a;
    }

struct GeometrySphere : Geometry
    {
    GeometrySphere(std::shared_ptr<Scene> s, unsigned N)
        {
        scene = s;
        check(fh_geometry_sphere_create(s->h, N, &h));
        }
    };
struct GeometryCylinder : Geometry
    {
    GeometryCylinder(std::shared_ptr<Scene> s, unsigned N)
        {
        scene = s;
        check(fh_geometry_cylinder_create(s->h, N, &h));
        }
    };
using farray = py::array_t<float, py::array::c_style | py::array::forcecast>;
struct GeometryPolygon : Geometry
    {
    GeometryPolygon(std::shared_ptr<Scene> s, farray verts, float rounding, unsigned N)
        {
        scene = s;
        auto i = verts.request();
        if (i.ndim != 2)
            throw std::runtime_error("vertices must be a 2-dimensional array");
        if (i.shape[1] != 2)
            throw std::runtime_error("vertices must be a Nvert by 2 array");
        check(fh_geometry_polygon_create(s->h, (const float*)i.ptr, (uint32_t)i.shape[0], rounding, N, &h));
        }
    float getRadius()
        {
        float r;
        check(fh_geometry_get_radius(h, &r));
        return r;
        }
    };
struct GeometryConvexPolyhedron : Geometry
    {
    GeometryConvexPolyhedron(std::shared_ptr<Scene> s, farray origins, farray normals, farray colors, unsigned N,
                             float r)
        {
        scene = s;
        auto o = origins.request(), n = normals.request(), c = colors.request();
        const char* names[3] = {"plane_origins", "plane_normals", "plane_colors"};
        py::buffer_info* all[3] = {&o, &n, &c};
        for (int k = 0; k < 3; k++)
            {
            if (all[k]->ndim != 2)
                throw std::runtime_error(std::string(names[k]) + " must be a 2-dimensional array");
            if (all[k]->shape[1] != 3)
                throw std::runtime_error(std::string(names[k]) + " must be a Nvert by 3 array");
            }
        if (n.shape[0] != o.shape[0])
            throw std::runtime_error("Number of vertices must match in origin and normal arrays");
        if (c.shape[0] != o.shape[0])
            throw std::runtime_error("Number of vertices must match in origin and color arrays");
        check(fh_geometry_convex_polyhedron_create(s->h, (const float*)o.ptr, (const float*)n.ptr,
                                                   (const float*)c.ptr, (uint32_t)o.shape[0], N, r, &h));
        }
    void setColorByFace(float f)
        {
        check(fh_geometry_set_color_by_face(h, f));
        }
    float getColorByFace()
        {
        float f;
        check(fh_geometry_get_color_by_face(h, &f));
        return f;
        }
    };
struct GeometryMesh : Geometry
    {
    GeometryMesh(std::shared_ptr<Scene> s, farray verts, unsigned N)
        {
        scene = s;
        auto i = verts.request();
        if (i.ndim != 2)
            throw std::runtime_error("vertices must be a 2-dimensional array");
        if (i.shape[1] != 3)
            throw std::runtime_error("vertices must be a Nvert by 3 array");
        if (i.shape[0] % 3 != 0)
            throw std::runtime_error("the number of triangle vertices must be a multiple of three.");
        check(fh_geometry_mesh_create(s->h, (const float*)i.ptr, (uint32_t)i.shape[0], N, &h));
        }
    };

struct Tracer
    {
    std::shared_ptr<Device> dev;
    fh_tracer h = nullptr;
    virtual ~Tracer()
        {
        fh_tracer_destroy(h);
        }
    void render(std::shared_ptr<Scene> s)
        {
        if (s->dev.get() != dev.get())
            throw std::runtime_error("Scene and Tracer devices do not match");
        s->push();
        check(fh_tracer_render(h, s->h));
        }
    void resize(unsigned w, unsigned hh)
        {
        check(fh_tracer_resize(h, w, hh));
        }
    void enableHighlightWarning(const fh_rgb& c)
        {
        check(fh_tracer_enable_highlight_warning(h, c));
        }
    void disableHighlightWarning()
        {
        check(fh_tracer_disable_highlight_warning(h));
        }
    void setSeed(unsigned s)
        {
        check(fh_tracer_set_seed(h, s));
        }
    unsigned getSeed()
        {
        uint32_t s;
        check(fh_tracer_get_seed(h, &s));
        return s;
        }
    };

static std::shared_ptr<Array> output_buffer(std::shared_ptr<Tracer> t, int which)
    {
    auto a = std::make_shared<Array>();
    a->owner = t;
    Tracer* tp = t.get();
    a->on_map = [tp, which]()
    {
        void* p;
        check(fh_tracer_map_output(tp->h, which, &p));
    };
    // map() copies the device image into the host staging buffer; the buffer
    // protocol exposes that buffer as the last map() left it
    a->info = [tp, which]()
    {
        void* p;
        check(fh_tracer_output_host(tp->h, which, &p));
        uint32_t w, hh;
        check(fh_tracer_get_size(tp->h, &w, &hh));
        if (which == FH_OUT_SRGB)
            return py::buffer_info(p, 1, py::format_descriptor<uint8_t>::format(), 3, {(py::ssize_t)hh, (py::ssize_t)w, (py::ssize_t)4},
                                   {(py::ssize_t)w * 4, (py::ssize_t)4, (py::ssize_t)1});
        return py::buffer_info(p, 4, py::format_descriptor<float>::format(), 3, {(py::ssize_t)hh, (py::ssize_t)w, (py::ssize_t)4},
                               {(py::ssize_t)w * 16, (py::ssize_t)16, (py::ssize_t)4});
    };
    a->layout = [tp, which]()
    {
        uint32_t w, hh;
        check(fh_tracer_get_size(tp->h, &w, &hh));
        return py::make_tuple(py::make_tuple(hh, w, 4), which == FH_OUT_SRGB ? "uint8" : "float32");
    };
    return a;
    }

struct TracerDirect : Tracer
    {
    TracerDirect(std::shared_ptr<Device> d, unsigned w, unsigned hh)
        {
        dev = d;
        check(fh_tracer_direct_create(d->h, w, hh, &h));
        }
    void setAntialiasingN(unsigned n)
        {
        check(fh_tracer_direct_set_aa_n(h, n));
        }
    unsigned getAntialiasingN()
        {
        uint32_t n;
        check(fh_tracer_direct_get_aa_n(h, &n));
        return n;
        }
    };

struct TracerPath : Tracer
    {
    TracerPath(std::shared_ptr<Device> d, unsigned w, unsigned hh, unsigned L)
        {
        dev = d;
        check(fh_tracer_path_create(d->h, w, hh, L, &h));
        }
    unsigned getNumSamples()
        {
        uint32_t n;
        check(fh_tracer_path_get_num_samples(h, &n));
        return n;
        }
    void reset()
        {
        check(fh_tracer_path_reset(h));
        }
    void setLightSamples(unsigned L)
        {
        check(fh_tracer_path_set_light_samples(h, L));
        }
    void renderSamples(std::shared_ptr<Scene> s, unsigned n)
        {
        if (s->dev.get() != dev.get())
            throw std::runtime_error("Scene and Tracer devices do not match");
        s->push();
        check(fh_tracer_render_samples(h, s->h, n));
        }
    };

// Render server: keeps the scene and tracer of each open request alive.
struct Server
    {
    std::shared_ptr<Device> dev;
    fh_server h = nullptr;
    std::map<uint64_t, std::pair<std::shared_ptr<Scene>, std::shared_ptr<TracerPath>>> open;
    Server(std::shared_ptr<Device> d, uint64_t budget, uint64_t target_paths) : dev(d)
        {
        check(fh_server_create(d->h, budget, target_paths, &h));
        }
    ~Server()
        {
        fh_server_destroy(h);
        }
    uint64_t submit(std::shared_ptr<Scene> s, std::shared_ptr<TracerPath> t, unsigned samples, unsigned L, double now)
        {
        if (s->dev.get() != dev.get() || t->dev.get() != dev.get())
            throw std::runtime_error("Scene and Tracer devices do not match");
        s->push();
        uint64_t id;
        check(fh_server_submit(h, s->h, t->h, samples, L, now, &id));
        open[id] = {s, t};
        return id;
        }
    py::tuple step(double now)
        {
        uint32_t a, q;
        check(fh_server_step(h, now, &a, &q));
        for (auto it = open.begin(); it != open.end();)
            {
            int st;
            check(fh_server_poll(h, it->first, &st, nullptr, nullptr));
            it = st >= 2 ? open.erase(it) : std::next(it);
            }
        return py::make_tuple(a, q);
        }
    py::tuple poll(uint64_t id)
        {
        int st;
        uint32_t done;
        double t[3];
        check(fh_server_poll(h, id, &st, &done, t));
        return py::make_tuple(st, done, py::make_tuple(t[0], t[1], t[2]));
        }
    void release(uint64_t id)
        {
        check(fh_server_release(h, id));
        }
    py::tuple memory()
        {
        uint64_t u, b;
        check(fh_server_memory(h, &u, &b));
        return py::make_tuple(u, b);
        }
    };

PYBIND11_MODULE(_gpu, m)
    {
    py::module_::import("fresnel_amd._common");

    m.def("get_num_available_devices",
          []()
          {
              int n = 0;
              check(fh_get_num_available_devices(&n));
              return n;
          });

    py::class_<Device, std::shared_ptr<Device>>(m, "Device")
        .def(py::init<const std::string&, int>())
        .def("describe", &Device::describe)
        .def_static("getAllGPUs", &Device::getAllGPUs)
        .def("statsEnable", [](Device& d, bool on) { check(fh_stats_enable(d.h, on ? 1 : 0)); })
        .def("countersEnable", [](Device& d, bool on) { check(fh_counters_enable(d.h, on ? 1 : 0)); })
        .def("statsLast",
             [](Device& d)
             {
                 char names[8192];
                 double ms[128];
                 uint32_t launches[128], n = 0;
                 check(fh_stats_last(d.h, names, sizeof(names), ms, launches, 128, &n));
                 py::dict out;
                 std::string all(names);
                 size_t pos = 0;
                 for (uint32_t k = 0; k < n && k < 128; k++)
                     {
                     size_t e = all.find('\n', pos);
                     std::string nm = all.substr(pos, e - pos);
                     pos = e + 1;
                     out[py::str(nm)] = py::make_tuple(ms[k], launches[k]);
                     }
                 return out;
             })
        .def("countersLast",
             [](Device& d)
             {
                 uint64_t c[16];
                 check(fh_counters_last(d.h, c));
                 return std::vector<uint64_t>(c, c + 16);
             })
        .def("debugCopy",
             [](Device& d, uint64_t bytes)
             {
                 float ms = 0;
                 check(fh_debug_copy(d.h, bytes, &ms));
                 return ms;
             })
        .def("tune",
             [](Device& d, const std::string& key, int64_t value)
             {
                 int64_t prev = 0;
                 check(fh_device_tune(d.h, key.c_str(), value, &prev));
                 return prev;
             })
        .def("debugSincos",
             [](Device& d, py::array_t<float, py::array::c_style | py::array::forcecast> x)
             {
                 auto i = x.request();
                 const py::ssize_t n = i.size;
                 py::array_t<float> s(n), c(n);
                 check(fh_debug_sincos(d.h, (const float*)i.ptr, (uint32_t)n, (float*)s.request().ptr,
                                       (float*)c.request().ptr));
                 return py::make_tuple(s, c);
             })
        .def("graphStats",
             [](Device& d)
             {
                 uint64_t c = 0, r = 0;
                 check(fh_stats_graphs(d.h, &c, &r));
                 return py::make_tuple(c, r);
             })
        .def("debugPhilox",
             [](Device& d, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> in)
             {
                 auto i = in.request();
                 if (i.ndim != 2 || i.shape[1] != 6)
                     throw std::runtime_error("input must be (n, 6) uint32");
                 py::array_t<uint32_t> out({(py::ssize_t)i.shape[0], (py::ssize_t)4});
                 check(fh_debug_philox(d.h, (const uint32_t*)i.ptr, (uint32_t)i.shape[0],
                                       (uint32_t*)out.request().ptr));
                 return out;
             });

    py::class_<Scene, std::shared_ptr<Scene>>(m, "Scene")
        .def(py::init<std::shared_ptr<Device>>())
        .def("setCamera", [](Scene& s, const fh_user_camera& c) { s.camera = c; })
        .def("getCamera", [](Scene& s) -> fh_user_camera& { return s.camera; }, py::return_value_policy::reference_internal)
        .def("getBackgroundColor", [](Scene& s) { return s.bg; })
        .def("setBackgroundColor", [](Scene& s, const fh_rgb& c) { s.bg = c; })
        .def("getBackgroundAlpha", [](Scene& s) { return s.bg_alpha; })
        .def("setBackgroundAlpha", [](Scene& s, float a) { s.bg_alpha = a; })
        .def("getLights", [](Scene& s) -> fh_lights& { return s.lights; }, py::return_value_policy::reference_internal)
        .def("setLights", [](Scene& s, const fh_lights& l) { s.lights = l; })
        .def("setRefit", [](Scene& s, bool on) { check(fh_scene_set_refit(s.h, on ? 1 : 0)); })
        .def("setRefitLimit", [](Scene& s, float limit) { check(fh_scene_set_refit_limit(s.h, limit)); })
        .def("bvhCost",
             [](Scene& s)
             {
                 float b, c;
                 check(fh_scene_bvh_cost(s.h, &b, &c));
                 return py::make_tuple(b, c);
             })
        .def("bvhInfo",
             [](Scene& s)
             {
                 uint64_t b, r;
                 uint32_t n;
                 check(fh_scene_bvh_info(s.h, &b, &r, &n));
                 return py::make_tuple(b, r, n);
             })
        .def("debugBVH",
             [](Scene& s)
             {
                 s.push();
                 uint64_t n = 0;
                 check(fh_debug_bvh(s.h, nullptr, 0, &n));
                 py::array_t<uint32_t> out({(py::ssize_t)n, (py::ssize_t)16});
                 check(fh_debug_bvh(s.h, out.request().ptr, n, &n));
                 return out;
             })
        .def("debugTrace",
             [](Scene& s, py::array_t<float, py::array::c_style | py::array::forcecast> rays, float tnear)
             {
                 auto r = rays.request();
                 if (r.ndim != 2 || r.shape[1] != 6)
                     throw std::runtime_error("rays must be (n, 6) float32");
                 const py::ssize_t n = r.shape[0];
                 py::array_t<uint32_t> gid(n), prim(n);
                 py::array_t<float> t(n);
                 s.push();
                 check(fh_debug_trace(s.h, (const float*)r.ptr, (uint32_t)n, tnear, (uint32_t*)gid.request().ptr,
                                      (uint32_t*)prim.request().ptr, (float*)t.request().ptr));
                 return py::make_tuple(gid, prim, t);
             });

    py::class_<Array, std::shared_ptr<Array>>(m, "Array", py::buffer_protocol())
        .def_buffer([](Array& a) { return a.info(); })
        .def("map", &Array::map)
        .def("unmap", &Array::unmap)
        .def("update_rows", &Array::update_rows)
        .def_property_readonly("layout", [](Array& a) { return a.layout(); });

    py::class_<Geometry, std::shared_ptr<Geometry>>(m, "Geometry")
        .def("getMaterial", &Geometry::getMaterial)
        .def("setMaterial", &Geometry::setMaterial)
        .def("getOutlineMaterial", &Geometry::getOutlineMaterial)
        .def("setOutlineMaterial", &Geometry::setOutlineMaterial)
        .def("getOutlineWidth", &Geometry::getOutlineWidth)
        .def("setOutlineWidth", &Geometry::setOutlineWidth)
        .def("disable", &Geometry::disable)
        .def("enable", &Geometry::enable)
        .def("remove", &Geometry::remove)
        .def("update", &Geometry::update);

    auto pos = [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_POSITION); };
    auto rad = [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_RADIUS); };
    auto col = [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_COLOR); };
    auto ori = [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_ORIENTATION); };

    py::class_<GeometrySphere, Geometry, std::shared_ptr<GeometrySphere>>(m, "GeometrySphere")
        .def(py::init<std::shared_ptr<Scene>, unsigned>())
        .def("getPositionBuffer", pos)
        .def("getRadiusBuffer", rad)
        .def("getColorBuffer", col);
    py::class_<GeometryCylinder, Geometry, std::shared_ptr<GeometryCylinder>>(m, "GeometryCylinder")
        .def(py::init<std::shared_ptr<Scene>, unsigned>())
        .def("getPointsBuffer", [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_POINTS); })
        .def("getRadiusBuffer", rad)
        .def("getColorBuffer", col);
    py::class_<GeometryPolygon, Geometry, std::shared_ptr<GeometryPolygon>>(m, "GeometryPolygon")
        .def(py::init<std::shared_ptr<Scene>, farray, float, unsigned>())
        .def("getPositionBuffer", pos)
        .def("getAngleBuffer", [](std::shared_ptr<Geometry> g) { return geometry_buffer(g, FH_BUF_ANGLE); })
        .def("getColorBuffer", col)
        .def("getRadius", &GeometryPolygon::getRadius);
    py::class_<GeometryConvexPolyhedron, Geometry, std::shared_ptr<GeometryConvexPolyhedron>>(m, "GeometryConvexPolyhedron")
        .def(py::init<std::shared_ptr<Scene>, farray, farray, farray, unsigned, float>())
        .def("getPositionBuffer", pos)
        .def("getOrientationBuffer", ori)
        .def("getColorBuffer", col)
        .def("setColorByFace", &GeometryConvexPolyhedron::setColorByFace)
        .def("getColorByFace", &GeometryConvexPolyhedron::getColorByFace);
    py::class_<GeometryMesh, Geometry, std::shared_ptr<GeometryMesh>>(m, "GeometryMesh")
        .def(py::init<std::shared_ptr<Scene>, farray, unsigned>())
        .def("getPositionBuffer", pos)
        .def("getOrientationBuffer", ori)
        .def("getColorBuffer", col);

    py::class_<Tracer, std::shared_ptr<Tracer>>(m, "Tracer")
        .def("render", &Tracer::render)
        .def("resize", &Tracer::resize)
        .def("getSRGBOutputBuffer", [](std::shared_ptr<Tracer> t) { return output_buffer(t, FH_OUT_SRGB); })
        .def("getLinearOutputBuffer", [](std::shared_ptr<Tracer> t) { return output_buffer(t, FH_OUT_LINEAR); })
        .def("histogram",
             [](std::shared_ptr<Tracer> t)
             {
                 py::array_t<uint32_t> out({(py::ssize_t)512, (py::ssize_t)4});
                 check(fh_tracer_histogram(t->h, out.mutable_data()));
                 return out;
             })
        .def("handle", [](std::shared_ptr<Tracer> t) { return (uintptr_t)t->h; })
        .def("outputAsync", [](std::shared_ptr<Tracer> t, int which) { check(fh_tracer_output_async(t->h, which)); })
        .def("outputReady",
             [](std::shared_ptr<Tracer> t, int which)
             {
                 int r = 0;
                 check(fh_tracer_output_ready(t->h, which, &r));
                 return r != 0;
             })
        .def("outputWait",
             [](std::shared_ptr<Tracer> t, int which)
             {
                 void* p = nullptr;
                 uint32_t w = 0, hh = 0;
                 check(fh_tracer_output_wait(t->h, which, &p, &w, &hh));
                 // a view of the tracer's pinned buffer (kept alive by the tracer)
                 py::object base = py::cast(t);
                 if (which == FH_OUT_SRGB)
                     return py::array(py::dtype::of<uint8_t>(), {(py::ssize_t)hh, (py::ssize_t)w, (py::ssize_t)4},
                                      {(py::ssize_t)w * 4, (py::ssize_t)4, (py::ssize_t)1}, p, base);
                 return py::array(py::dtype::of<float>(), {(py::ssize_t)hh, (py::ssize_t)w, (py::ssize_t)4},
                                  {(py::ssize_t)w * 16, (py::ssize_t)16, (py::ssize_t)4}, p, base);
             })
        .def("enableHighlightWarning", &Tracer::enableHighlightWarning)
        .def("disableHighlightWarning", &Tracer::disableHighlightWarning)
        .def("getSeed", &Tracer::getSeed)
        .def("setSeed", &Tracer::setSeed);
    py::class_<TracerDirect, Tracer, std::shared_ptr<TracerDirect>>(m, "TracerDirect")
        .def(py::init<std::shared_ptr<Device>, unsigned, unsigned>())
        .def("setAntialiasingN", &TracerDirect::setAntialiasingN)
        .def("getAntialiasingN", &TracerDirect::getAntialiasingN);
    py::class_<TracerPath, Tracer, std::shared_ptr<TracerPath>>(m, "TracerPath")
        .def(py::init<std::shared_ptr<Device>, unsigned, unsigned, unsigned>())
        .def("getNumSamples", &TracerPath::getNumSamples)
        .def("reset", &TracerPath::reset)
        .def("setLightSamples", &TracerPath::setLightSamples)
        .def("renderSamples", &TracerPath::renderSamples)
        .def("debugPrimaryRender",
             [](TracerPath& t, std::shared_ptr<Scene> s, unsigned sample)
             {
                 uint32_t w, hh;
                 check(fh_tracer_get_size(t.h, &w, &hh));
                 const py::ssize_t n = (py::ssize_t)w * hh;
                 py::array_t<uint32_t> gid(n), prim(n);
                 py::array_t<float> tt(n);
                 s->push();
                 check(fh_debug_primary_render(t.h, s->h, sample, (uint32_t*)gid.request().ptr,
                                               (uint32_t*)prim.request().ptr, (float*)tt.request().ptr));
                 return py::make_tuple(gid, prim, tt);
             })
        .def("debugPrimary",
             [](TracerPath& t, std::shared_ptr<Scene> s, unsigned sample)
             {
                 uint32_t w, hh;
                 check(fh_tracer_get_size(t.h, &w, &hh));
                 const py::ssize_t n = (py::ssize_t)w * hh;
                 py::array_t<float> rays({n, (py::ssize_t)6});
                 py::array_t<uint32_t> gid(n), prim(n);
                 py::array_t<float> tt(n);
                 s->push();
                 check(fh_debug_primary(t.h, s->h, sample, (float*)rays.request().ptr, (uint32_t*)gid.request().ptr,
                                        (uint32_t*)prim.request().ptr, (float*)tt.request().ptr));
                 return py::make_tuple(rays, gid, prim, tt);
             });

    py::class_<Server, std::shared_ptr<Server>>(m, "Server")
        .def(py::init<std::shared_ptr<Device>, uint64_t, uint64_t>())
        .def("submit", &Server::submit, py::arg("scene"), py::arg("tracer"), py::arg("samples"),
             py::arg("light_samples") = 1, py::arg("now") = -1.0)
        .def("step", &Server::step, py::arg("now") = -1.0)
        .def("poll", &Server::poll)
        .def("release", &Server::release)
        .def("memory", &Server::memory);
    m.def("batch_render",
          [](std::vector<std::shared_ptr<Scene>> scenes, std::vector<std::shared_ptr<TracerPath>> tracers,
             unsigned samples)
          {
              if (scenes.size() != tracers.size())
                  throw std::runtime_error("scenes and tracers must have the same length");
              std::vector<fh_scene> sh;
              std::vector<fh_tracer> th;
              for (size_t k = 0; k < scenes.size(); k++)
                  {
                  if (scenes[k]->dev.get() != tracers[k]->dev.get())
                      throw std::runtime_error("Scene and Tracer devices do not match");
                  scenes[k]->push();
                  sh.push_back(scenes[k]->h);
                  th.push_back(tracers[k]->h);
                  }
              check(fh_batch_render(sh.data(), th.data(), (uint32_t)sh.size(), samples));
          });
    }

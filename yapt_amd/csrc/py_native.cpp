// pybind11 binding of the C ABI (include/yapt_pt.h): the compiled layer under the py_yapt surface.
//
// The reference's Python surface is the pybind11 module py_yapt (yapt_harness.py:48-53), whose
// renderer drives a PathIntegratorSubStage. Here the integrator is the C ABI's handle, and this
// module binds it for Python: `Integrator` (ypt_create .. ypt_destroy) and `ResidentScene`
// (ypt_scene_create). Scene descriptions come from an object with the arrays that
// yapt_amd._lib.SceneArrays holds (numpy arrays, and ctypes arrays of the C records, read through
// the buffer protocol); frame constants and post-processing parameters are buffers holding one
// ypt_frame_desc / ypt_post_params. Results are numpy arrays. Errors raise RuntimeError with the
// library's message; the GIL is released around every call that waits on the GPU.
//
// Built by yapt_amd/build.py with g++ against libyapt_pt.so (host code only). Import it after
// yapt_amd._lib.load(), which loads the library with the HIP runtime that torch shares.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/yapt_pt.h"

namespace py = pybind11;

namespace {

void check(int rc) {
  if (rc != YPT_OK) throw std::runtime_error("yapt_pt error " + std::to_string(rc) + ": " + ypt_last_error());
}

// the data pointer of a buffer whose elements are `item` bytes each (count elements at least); None -> null
const void* buffer_ptr(const py::object& o, size_t item, size_t count, const char* what) {
  if (o.is_none()) {
    if (count) throw std::invalid_argument(std::string(what) + " is missing");
    return nullptr;
  }
  py::buffer_info bi = py::buffer(o).request();
  size_t bytes = (size_t)bi.size * (size_t)bi.itemsize;
  if (bytes < item * count) throw std::invalid_argument(std::string(what) + " holds fewer bytes than its count needs");
  return bi.ptr;
}

// the data of a C-contiguous numpy array of T with at least `count` elements; None -> null
template <class T>
const T* array_ptr(const py::object& o, size_t count, const char* what) {
  if (o.is_none()) {
    if (count) throw std::invalid_argument(std::string(what) + " is missing");
    return nullptr;
  }
  if (!py::isinstance<py::array_t<T>>(o)) throw std::invalid_argument(std::string(what) + ": wrong element type");
  py::array a = py::reinterpret_borrow<py::array>(o);
  if (!(a.flags() & py::array::c_style)) throw std::invalid_argument(std::string(what) + " is not C-contiguous");
  if ((size_t)a.size() < count) throw std::invalid_argument(std::string(what) + " holds fewer elements than its count needs");
  return static_cast<const T*>(a.data());
}

// an optional per-vertex array: None -> null, else as array_ptr
template <class T>
const T* optional_array_ptr(const py::object& o, size_t count, const char* what) {
  return o.is_none() ? nullptr : array_ptr<T>(o, count, what);
}

size_t array_size(const py::object& o) { return (size_t)py::reinterpret_borrow<py::array>(o).size(); }

py::object attr_or_none(const py::object& o, const char* name) {
  if (py::hasattr(o, name)) return o.attr(name);
  return py::none();
}

// A scene description over the arrays of a SceneArrays-like object. Texture and environment records
// carry pointers into that object's texel arrays, so the object is kept alive with the description.
struct SceneDescView {
  ypt_scene_desc d;
  py::object keep;
  explicit SceneDescView(const py::object& a) : keep(a) {
    std::memset(&d, 0, sizeof(d));
    // the object's own arrays (C-contiguous, of the element type the ABI takes), so pointers stay valid
    d.vertex_count = (uint32_t)(array_size(a.attr("positions")) / 3);
    d.triangle_count = (uint32_t)(array_size(a.attr("indices")) / 3);
    d.positions = array_ptr<float>(a.attr("positions"), 3 * (size_t)d.vertex_count, "positions");
    d.indices = array_ptr<uint32_t>(a.attr("indices"), 3 * (size_t)d.triangle_count, "indices");
    d.normals = optional_array_ptr<float>(attr_or_none(a, "normals"), 3 * (size_t)d.vertex_count, "normals");
    d.tangents = optional_array_ptr<float>(attr_or_none(a, "tangents"), 4 * (size_t)d.vertex_count, "tangents");
    d.uvs = optional_array_ptr<float>(attr_or_none(a, "uvs"), 2 * (size_t)d.vertex_count, "uvs");
    d.instance_count = (uint32_t)array_size(a.attr("inst_first"));
    d.instance_first_triangle = array_ptr<uint32_t>(a.attr("inst_first"), d.instance_count, "inst_first");
    d.instance_triangle_count = array_ptr<uint32_t>(a.attr("inst_count"), d.instance_count, "inst_count");
    d.instance_material = array_ptr<uint32_t>(a.attr("inst_material"), d.instance_count, "inst_material");
    py::object mats = a.attr("materials");
    d.material_count = (uint32_t)py::len(mats);
    d.materials = static_cast<const ypt_material*>(buffer_ptr(mats, sizeof(ypt_material), d.material_count, "materials"));
    d.punctual_light_count = (uint32_t)py::len(a.attr("lights"));
    d.punctual_lights = d.punctual_light_count
                            ? static_cast<const ypt_punctual_light*>(buffer_ptr(a.attr("light_array"), sizeof(ypt_punctual_light),
                                                                                d.punctual_light_count, "light_array"))
                            : nullptr;
    d.texture_count = (uint32_t)py::len(a.attr("texel_arrays"));
    d.textures = d.texture_count ? static_cast<const ypt_texture*>(buffer_ptr(a.attr("texture_array"), sizeof(ypt_texture),
                                                                              d.texture_count, "texture_array"))
                                 : nullptr;
    py::object env = attr_or_none(a, "env");
    d.environment = env.is_none() ? nullptr
                                  : static_cast<const ypt_environment*>(buffer_ptr(env, sizeof(ypt_environment), 1, "env"));
  }
};

struct ResidentScene {
  ypt_scene* s = nullptr;
  uint64_t bytes = 0;
  uint32_t triangles = 0;
  ResidentScene(const py::object& arrays, int device) {
    SceneDescView v(arrays);
    {
      py::gil_scoped_release nogil;
      check(ypt_scene_create(&v.d, device, &s));
    }
    check(ypt_scene_info(s, &bytes, &triangles));
  }
  ~ResidentScene() { close(); }
  void close() {
    if (s) ypt_scene_destroy(s);
    s = nullptr;
  }
};

struct Integrator {
  ypt_integrator* h = nullptr;
  uint32_t width = 0, height = 0;
  py::object scene;  // what set_scene received, kept alive with the handle
  Integrator(int type, int device) { check(ypt_create(type, device, &h)); }
  ~Integrator() { close(); }
  void close() {
    if (h) ypt_destroy(h);
    h = nullptr;
  }
  ypt_integrator* handle() const {
    if (!h) throw std::runtime_error("the integrator is closed");
    return h;
  }
  void set_option(const std::string& name, int64_t value) { check(ypt_set_option(handle(), name.c_str(), value)); }
  void set_scene(const py::object& sc) {
    if (py::isinstance<ResidentScene>(sc)) {
      check(ypt_set_resident_scene(handle(), sc.cast<ResidentScene&>().s));
    } else {
      SceneDescView v(sc);
      py::gil_scoped_release nogil;
      check(ypt_set_scene(handle(), &v.d));
    }
    scene = sc;
  }
  void refit_scene(const py::object& sc) {
    SceneDescView v(sc);
    {
      py::gil_scoped_release nogil;
      check(ypt_refit_scene(handle(), &v.d));
    }
    scene = sc;
  }
  py::tuple bvh_counts() {
    uint32_t b = 0, r = 0;
    check(ypt_scene_bvh_counts(handle(), &b, &r));
    return py::make_tuple(b, r);
  }
  void set_frame(const py::object& frame) {
    const ypt_frame_desc* f = static_cast<const ypt_frame_desc*>(buffer_ptr(frame, sizeof(ypt_frame_desc), 1, "frame"));
    check(ypt_set_frame(handle(), f));
    width = f->width;
    height = f->height;
  }
  void clear_film() { check(ypt_clear_film(handle())); }
  void render(uint32_t begin, uint32_t count) {
    py::gil_scoped_release nogil;
    check(ypt_render(handle(), begin, count, nullptr));
  }
  void synchronize() {
    py::gil_scoped_release nogil;
    check(ypt_synchronize(handle()));
  }
  py::array_t<float> image(int channels, int (*get)(ypt_integrator*, float*)) {
    py::array_t<float> out({(py::ssize_t)height, (py::ssize_t)width, (py::ssize_t)channels});
    float* p = out.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(get(handle(), p));
    }
    return out;
  }
  py::array_t<float> film() {
    return image(4, [](ypt_integrator* x, float* p) { return ypt_get_film(x, p, nullptr); });
  }
  py::array_t<float> last_samples() { return image(4, ypt_get_last_samples); }
  py::array_t<float> first_hit_features() { return image(12, ypt_get_first_hit_features); }
  py::object final_color(const py::object& params, bool linear) {
    ypt_post_params dflt;
    const ypt_post_params* p;
    if (params.is_none()) {
      ypt_post_params_default(&dflt);
      p = &dflt;
    } else {
      p = static_cast<const ypt_post_params*>(buffer_ptr(params, sizeof(ypt_post_params), 1, "params"));
    }
    py::array_t<uint8_t> rgba({(py::ssize_t)height, (py::ssize_t)width, (py::ssize_t)4});
    py::array_t<float> lin;
    float* linp = nullptr;
    if (linear) {
      lin = py::array_t<float>({(py::ssize_t)height, (py::ssize_t)width, (py::ssize_t)4});
      linp = lin.mutable_data();
    }
    uint8_t* rp = rgba.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(ypt_final_color(handle(), p, rp, linp));
    }
    if (linear) return py::make_tuple(rgba, lin);
    return rgba;
  }
  py::array_t<double> stats() {
    py::array_t<double> out(16);
    check(ypt_get_stats(handle(), out.mutable_data()));
    return out;
  }
  py::tuple trace_closest(py::array_t<float, py::array::c_style | py::array::forcecast> rays) {
    if (rays.size() % 8) throw std::invalid_argument("rays: 8 floats per ray (origin, tmin, direction, tmax)");
    uint32_t n = (uint32_t)(rays.size() / 8);
    py::array_t<float> t(n), b({(py::ssize_t)n, (py::ssize_t)2});
    py::array_t<uint32_t> tri(n);
    const float* rp = rays.data();
    float *tp = t.mutable_data(), *bp = b.mutable_data();
    uint32_t* ip = tri.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(ypt_trace_closest(handle(), rp, n, tp, ip, bp));
    }
    return py::make_tuple(t, tri, b);
  }
};

}  // namespace

PYBIND11_MODULE(_yapt_native, m) {
  m.doc() = "pybind11 binding of libyapt_pt's C ABI: the integrator handle and resident scenes";
  m.attr("YPT_INTEGRATOR_BPT") = YPT_INTEGRATOR_BPT;
  m.attr("YPT_INTEGRATOR_NEE") = YPT_INTEGRATOR_NEE;
  m.attr("YPT_INTEGRATOR_BDPT") = YPT_INTEGRATOR_BDPT;
  m.attr("FRAME_DESC_BYTES") = sizeof(ypt_frame_desc);
  m.attr("MATERIAL_BYTES") = sizeof(ypt_material);
  m.def("version", []() { return std::string(ypt_version()); });
  m.def("scene_desc_counts", [](const py::object& arrays) {  // what set_scene would pass (host only)
    SceneDescView v(arrays);
    py::dict out;
    out["vertices"] = v.d.vertex_count;
    out["triangles"] = v.d.triangle_count;
    out["instances"] = v.d.instance_count;
    out["materials"] = v.d.material_count;
    out["punctual_lights"] = v.d.punctual_light_count;
    out["textures"] = v.d.texture_count;
    out["normals"] = v.d.normals != nullptr;
    out["tangents"] = v.d.tangents != nullptr;
    out["uvs"] = v.d.uvs != nullptr;
    out["environment"] = v.d.environment != nullptr;
    return out;
  });
  m.def("sobol_point", [](uint32_t index) {
    py::array_t<uint32_t> out(256);
    check(ypt_sobol_point(index, out.mutable_data()));
    return out;
  });
  py::class_<ResidentScene>(m, "ResidentScene", "a scene built once on a device (ypt_scene_create)")
      .def(py::init<const py::object&, int>(), py::arg("arrays"), py::arg("device") = 0)
      .def_readonly("device_bytes", &ResidentScene::bytes)
      .def_readonly("triangle_count", &ResidentScene::triangles)
      .def("close", &ResidentScene::close);
  py::class_<Integrator>(m, "Integrator", "one path integrator on one GPU (PathIntegratorSubStage over ypt_*)")
      .def(py::init<int, int>(), py::arg("integrator_type") = YPT_INTEGRATOR_NEE, py::arg("device") = 0)
      .def_readonly("width", &Integrator::width)
      .def_readonly("height", &Integrator::height)
      .def("close", &Integrator::close)
      .def("set_option", &Integrator::set_option)
      .def("set_scene", &Integrator::set_scene)
      .def("set_frame", &Integrator::set_frame)
      .def("refit_scene", &Integrator::refit_scene, "moved vertex data for the same triangles: BVH refit, no rebuild")
      .def("bvh_counts", &Integrator::bvh_counts)
      .def("clear_film", &Integrator::clear_film)
      .def("render", &Integrator::render, py::arg("sample_begin"), py::arg("sample_count"))
      .def("synchronize", &Integrator::synchronize)
      .def("film", &Integrator::film)
      .def("last_samples", &Integrator::last_samples)
      .def("first_hit_features", &Integrator::first_hit_features)
      .def("final_color", &Integrator::final_color, py::arg("params") = py::none(), py::arg("linear") = false)
      .def("stats", &Integrator::stats)
      .def("trace_closest", &Integrator::trace_closest);
}

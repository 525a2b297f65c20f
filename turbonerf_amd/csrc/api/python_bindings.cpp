// PyTurboNeRF: pybind11 surface mirroring the reference module
// (src/api/python_bindings.cu:56-694) for the training/rendering path.
#include <pybind11/functional.h>
#include <cmath>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "autotune.hpp"
#include "controllers.hpp"
#include "file_manager.hpp"
#include "renderer.hpp"

namespace py = pybind11;
using namespace turbo;

namespace {

int2v to_int2(py::handle h) {
    auto s = h.cast<py::sequence>();
    if (py::len(s) != 2) throw std::runtime_error("expected a pair");
    return int2v{s[0].cast<int>(), s[1].cast<int>()};
}

float2v to_float2(py::handle h) {
    auto s = h.cast<py::sequence>();
    if (py::len(s) != 2) throw std::runtime_error("expected a pair");
    return float2v{s[0].cast<float>(), s[1].cast<float>()};
}

// host copy of the first n elements of a device buffer (n = 0: all)
py::dict autotune_dict(const AutotuneResult& r) {
    py::dict d;
    d["picks"] = r.picks;
    d["trials"] = r.trials;
    d["ns_per_sample_before"] = r.ns_per_sample_before;
    d["ns_per_sample_after"] = r.ns_per_sample_after;
    return d;
}

template <typename T>
py::array_t<T> to_numpy(const DeviceBuffer<T>& b, size_t n = 0) {
    if (n == 0 || n > b.size()) n = b.size();
    py::array_t<T> a(n);
    if (n) TN_HIP(hipMemcpy(a.mutable_data(), b.data(), n * sizeof(T), hipMemcpyDeviceToHost));
    return a;
}

// device buffer <- host array of exactly its size
template <typename T>
void from_numpy(DeviceBuffer<T>& b, py::array_t<T, py::array::c_style | py::array::forcecast> a) {
    if ((size_t)a.size() != b.size()) throw std::runtime_error("size mismatch");
    TN_HIP(hipMemcpy(b.data(), a.data(), b.bytes(), hipMemcpyHostToDevice));
}

// fp32 host copy of a precision-typed (fp16 or fp32) byte buffer
py::array_t<float> typed_to_float(const DeviceBuffer<uint8_t>& b, int precision) {
    const size_t esz = precision == 1 ? 2 : 4, n = b.size() / esz;
    py::array_t<float> a(n);
    if (precision == 1) {
        std::vector<uint16_t> h(n);
        TN_HIP(hipMemcpy(h.data(), b.data(), n * 2, hipMemcpyDeviceToHost));
        float* o = a.mutable_data();
        for (size_t i = 0; i < n; ++i) {
            const uint32_t x = h[i], sign = (x >> 15) & 1, e = (x >> 10) & 31, m = x & 1023;
            float v = e == 0 ? std::ldexp((float)m, -24) : e == 31 ? (m ? NAN : INFINITY)
                                                                 : std::ldexp((float)(m | 1024), (int)e - 25);
            o[i] = sign ? -v : v;
        }
    } else if (n) {
        TN_HIP(hipMemcpy(a.mutable_data(), b.data(), n * 4, hipMemcpyDeviceToHost));
    }
    return a;
}

Transform4f to_transform(py::array_t<float, py::array::c_style | py::array::forcecast> a) {
    if (a.ndim() != 2 || a.shape(1) != 4 || (a.shape(0) != 3 && a.shape(0) != 4))
        throw std::runtime_error("Invalid shape for Transform4f");
    return Transform4f::from_array(a.data());
}

}  // namespace

PYBIND11_MODULE(PyTurboNeRF, m) {
    m.doc() = "TurboNeRF (MI355X) Python bindings";
    m.attr("__version__") = "0.0.19";
    m.attr("__build__") = 1;
    m.def("teardown", []() { DeviceManager::teardown(); });

    py::class_<RuntimeVersion>(m, "RuntimeVersion")
        .def_readonly("major", &RuntimeVersion::major)
        .def_readonly("minor", &RuntimeVersion::minor)
        .def_readonly("patch", &RuntimeVersion::patch)
        .def("to_string", &RuntimeVersion::to_string)
        .def("__str__", &RuntimeVersion::to_string)
        .def("__ge__", &RuntimeVersion::operator>=);

    // reference src/api/python_bindings.cu:685-692 (cuda_archs becomes gpu_archs)
    py::class_<RuntimeManager>(m, "RuntimeManager")
        .def(py::init<>())
        .def_readonly("gpu_archs", &RuntimeManager::gpu_archs)
        .def_static("required_runtime_version", &RuntimeManager::required_runtime_version)
        .def_static("compiled_runtime_version", &RuntimeManager::compiled_runtime_version)
        .def_static("current_runtime_version", &RuntimeManager::current_runtime_version)
        .def_static("is_driver_version_sufficient", &RuntimeManager::is_driver_version_sufficient)
        .def_static("get_device_architecture", &RuntimeManager::get_device_architecture, py::arg("device_id"))
        .def("check_runtime", &RuntimeManager::check_runtime);

    py::class_<Transform4f>(m, "Transform4f", py::buffer_protocol())
        .def(py::init([](py::array_t<float, py::array::c_style | py::array::forcecast> a) { return to_transform(a); }),
             py::arg("matrix"))
        .def_static("Identity", &Transform4f::Identity)
        .def("from_nerf", &Transform4f::from_nerf)
        .def("to_nerf", &Transform4f::to_nerf)
        .def("inverse", &Transform4f::inverse)
        .def("to_matrix", [](const Transform4f& t) {
            auto a = t.to_array();
            py::array_t<float> out({3, 4});
            std::memcpy(out.mutable_data(), a.data(), sizeof(a));
            return out;
        })
        .def(py::self == py::self);

    py::class_<DistortionParams>(m, "DistortionParams")
        .def(py::init([](float k1, float k2, float k3, float p1, float p2) {
                 return DistortionParams{k1, k2, k3, p1, p2};
             }),
             py::arg("k1") = 0.0f, py::arg("k2") = 0.0f, py::arg("k3") = 0.0f, py::arg("p1") = 0.0f,
             py::arg("p2") = 0.0f)
        .def_readwrite("k1", &DistortionParams::k1)
        .def_readwrite("k2", &DistortionParams::k2)
        .def_readwrite("k3", &DistortionParams::k3)
        .def_readwrite("p1", &DistortionParams::p1)
        .def_readwrite("p2", &DistortionParams::p2);

    py::class_<Camera>(m, "Camera")
        .def(py::init([](py::object res, float near, float far, py::object focal, py::object pp, py::object shift,
                         const Transform4f& t, const DistortionParams& dp) {
                 return Camera(to_int2(res), near, far, to_float2(focal), to_float2(pp), to_float2(shift), t, dp);
             }),
             py::arg("resolution"), py::arg("near"), py::arg("far"), py::arg("focal_length"),
             py::arg("principal_point"), py::arg("shift"), py::arg("transform"),
             py::arg("dist_params") = DistortionParams())
        .def_property("resolution", [](const Camera& c) { return py::make_tuple(c.resolution.x, c.resolution.y); },
                      [](Camera& c, py::object v) { c.resolution = to_int2(v); })
        .def_readwrite("near", &Camera::near)
        .def_readwrite("far", &Camera::far)
        .def_property("focal_length", [](const Camera& c) { return py::make_tuple(c.focal_length.x, c.focal_length.y); },
                      [](Camera& c, py::object v) { c.focal_length = to_float2(v); })
        .def_property("principal_point",
                      [](const Camera& c) { return py::make_tuple(c.principal_point.x, c.principal_point.y); },
                      [](Camera& c, py::object v) { c.principal_point = to_float2(v); })
        .def_property("shift", [](const Camera& c) { return py::make_tuple(c.shift.x, c.shift.y); },
                      [](Camera& c, py::object v) { c.shift = to_float2(v); })
        .def_readwrite("transform", &Camera::transform)
        .def_readwrite("dist_params", &Camera::dist_params)
        .def_readwrite("show_image_planes", &Camera::show_image_planes)
        .def(py::self == py::self)
        .def(py::self != py::self);

    py::class_<BoundingBox>(m, "BoundingBox")
        .def(py::init<>())
        .def(py::init<float>(), py::arg("size"))
        .def(py::init([](float x0, float y0, float z0, float x1, float y1, float z1) {
                 BoundingBox b;
                 b.min_x = x0; b.min_y = y0; b.min_z = z0;
                 b.max_x = x1; b.max_y = y1; b.max_z = z1;
                 return b;
             }),
             py::arg("min_x"), py::arg("min_y"), py::arg("min_z"), py::arg("max_x"), py::arg("max_y"),
             py::arg("max_z"))
        .def("size", &BoundingBox::size)
        .def(py::self == py::self)
        .def_readwrite("min_x", &BoundingBox::min_x)
        .def_readwrite("min_y", &BoundingBox::min_y)
        .def_readwrite("min_z", &BoundingBox::min_z)
        .def_readwrite("max_x", &BoundingBox::max_x)
        .def_readwrite("max_y", &BoundingBox::max_y)
        .def_readwrite("max_z", &BoundingBox::max_z);

    py::class_<TrainingImage>(m, "TrainingImage")
        .def(py::init([](const std::string& path, py::object dims) { return TrainingImage{path, to_int2(dims)}; }),
             py::arg("file_path"), py::arg("dimensions"))
        .def_readonly("file_path", &TrainingImage::file_path)
        .def_property_readonly("dimensions", [](const TrainingImage& t) {
            return py::make_tuple(t.dimensions.x, t.dimensions.y);
        });

    py::class_<Dataset>(m, "Dataset")
        .def(py::init<const std::string&>(), py::arg("file_path"))
        .def(py::init([](const BoundingBox& bbox, const std::vector<Camera>& cams,
                         py::array_t<uint8_t, py::array::c_style | py::array::forcecast> images) {
                 if (images.ndim() != 4 || images.shape(3) != 4 || (size_t)images.shape(0) != cams.size())
                     throw std::runtime_error("images must be (n_cameras, height, width, 4) uint8");
                 return Dataset(bbox, cams, images.data(), (int)images.shape(2), (int)images.shape(1));
             }),
             py::arg("bounding_box"), py::arg("cameras"), py::arg("images"))
        .def("load_transforms", &Dataset::load_transforms)
        .def("copy", &Dataset::copy)
        .def("to_json", [](const Dataset& d) { return py::module_::import("json").attr("loads")(d.to_json()); })
        .def_property_readonly("images", &Dataset::images)
        .def("set_camera_at", [](Dataset& ds, int i, const Camera& c) { ds.cameras.at(i) = c; })
        .def_readwrite("file_path", &Dataset::file_path)
        .def_readwrite("cameras", &Dataset::cameras)
        .def_readwrite("bounding_box", &Dataset::bounding_box)
        .def_property_readonly("image_dimensions",
                               [](const Dataset& d) { return py::make_tuple(d.image_dimensions.x, d.image_dimensions.y); })
        .def_property_readonly("n_images", &Dataset::n_images)
        .def_readonly("image_paths", &Dataset::image_paths)
        .def("is_loaded_cpu", &Dataset::is_loaded_cpu)
        .def("load_images", &Dataset::load_images, "Decode the image files (file-backed datasets).")
        .def("unload_images", &Dataset::unload_images)
        .def_property_readonly("image_data",
                               [](const Dataset& d) {
                                   if (!d.is_loaded_cpu()) throw std::runtime_error("images are not loaded");
                                   return py::array_t<uint8_t>({(py::ssize_t)d.n_images(),
                                                                (py::ssize_t)d.image_dimensions.y,
                                                                (py::ssize_t)d.image_dimensions.x, (py::ssize_t)4},
                                                               d.image_data.data());
                               })
        .def_readonly("contains_multiple_image_dims", &Dataset::contains_multiple_image_dims);

    // snapshots (reference python_bindings.cu:647-664)
    py::class_<FileManager>(m, "FileManager")
        .def_static(
            "load", [](NeRFProxy* nerf, const std::string& path) { return FileManager::load(nerf, path); },
            py::arg("nerf"), py::arg("path"),
            "Load a snapshot; returns True when it held optimiser state (restored with the training step).")
        .def_static(
            "save",
            [](NeRFProxy* nerf, const std::string& path, bool opt) { FileManager::save(nerf, path, opt); },
            py::arg("nerf"), py::arg("path"), py::arg("include_optimizer_state") = false);

    // UpdatableProperty<T> wrappers (reference python_bindings.cu:43-53,257-258)
    py::class_<UpdatableProperty<BoundingBox>>(m, "UpdatableBoundingBox")
        .def(py::init<>())
        .def("get", &UpdatableProperty<BoundingBox>::get)
        .def("set", &UpdatableProperty<BoundingBox>::set)
        .def("set_dirty", &UpdatableProperty<BoundingBox>::set_dirty)
        .def("is_dirty", &UpdatableProperty<BoundingBox>::is_dirty);
    py::class_<UpdatableProperty<Transform4f>>(m, "UpdatableTransform4f")
        .def(py::init<>())
        .def("get", &UpdatableProperty<Transform4f>::get)
        .def("set", &UpdatableProperty<Transform4f>::set)
        .def("set_dirty", &UpdatableProperty<Transform4f>::set_dirty)
        .def("is_dirty", &UpdatableProperty<Transform4f>::is_dirty);

    py::class_<NeRFProxy>(m, "NeRF")
        .def("attach_dataset", &NeRFProxy::attach_dataset)
        .def("detach_dataset", &NeRFProxy::detach_dataset)
        .def("is_dirty", &NeRFProxy::is_dirty)
        .def("can_train", &NeRFProxy::can_train)
        .def("is_image_data_loaded", &NeRFProxy::is_image_data_loaded)
        .def_property("render_bbox", [](const NeRFProxy& p) { return p.render_bbox.get(); },
                      [](NeRFProxy& p, const BoundingBox& b) { p.render_bbox.set(b); })
        .def_property("training_bbox", [](const NeRFProxy& p) { return p.training_bbox.get(); },
                      [](NeRFProxy& p, const BoundingBox& b) { p.training_bbox.set(b); })
        .def_property("transform", [](const NeRFProxy& p) { return p.transform.get(); },
                      [](NeRFProxy& p, const Transform4f& t) { p.transform.set(t); })
        .def_readwrite("is_visible", &NeRFProxy::is_visible)
        .def_readwrite("is_dataset_dirty", &NeRFProxy::is_dataset_dirty)
        .def_property_readonly("dataset", [](const NeRFProxy& p) { return p.dataset; })
        .def_readonly("can_render", &NeRFProxy::can_render)
        .def_readonly("training_step", &NeRFProxy::training_step)
        .def_readonly("id", &NeRFProxy::id)
        .def_property_readonly("n_appearances",
                               [](const NeRFProxy& p) { return p.nerfs.empty() ? 0u : p.nerfs[0].params.n_appearances; })
        .def(
            "set_analytic_field",
            [](NeRFProxy& p, py::array_t<float, py::array::c_style | py::array::forcecast> blobs, float aabb,
               float threshold) {
                if (blobs.ndim() != 2 || blobs.shape(1) != 8) throw std::runtime_error("blobs must be (n, 8)");
                set_analytic_field(&p, blobs.data(), (uint32_t)blobs.shape(0), aabb, threshold);
            },
            py::arg("blobs"), py::arg("aabb"), py::arg("threshold") = 0.01f,
            "Synthetic scene: render a Gaussian-blob field (rows cx, cy, cz, s, p, r, g, b) instead of a network.")
        .def("get_occupancy_bitfield", [](NeRFProxy& p) {
            DeviceManager::synchronize();
            const auto& og = p.nerfs[0].occupancy_grid;
            py::array_t<uint8_t> a(og.bitfield.size());
            TN_HIP(hipMemcpy(a.mutable_data(), og.bitfield.data(), og.bitfield.size(), hipMemcpyDeviceToHost));
            return a;
        });

    py::class_<NeRFManager>(m, "NeRFManager")
        .def(py::init<>())
        // the returned NeRF keeps its manager (the owner of its storage) alive
        .def("create", &NeRFManager::create, py::return_value_policy::reference_internal)
        .def("clone", &NeRFManager::clone, py::arg("nerf"), py::return_value_policy::reference)
        .def("destroy", &NeRFManager::destroy, py::arg("nerf"))
        .def("get_nerfs", &NeRFManager::get_proxies, py::return_value_policy::reference);

    py::class_<NetworkConfig>(m, "NetworkConfig")
        .def(py::init<>())
        .def_readwrite("precision", &NetworkConfig::precision)
        .def_readwrite("n_levels", &NetworkConfig::n_levels)
        .def_readwrite("log2_hashmap_size", &NetworkConfig::log2_hashmap_size)
        .def_readwrite("base_resolution", &NetworkConfig::base_resolution)
        .def_readwrite("finest_resolution_per_unit", &NetworkConfig::finest_resolution_per_unit)
        .def_readwrite("hash_init_scale", &NetworkConfig::hash_init_scale)
        .def_readwrite("seed", &NetworkConfig::seed)
        .def_readwrite("learning_rate", &NetworkConfig::learning_rate)
        .def_readwrite("learning_rate_decay", &NetworkConfig::learning_rate_decay)
        .def_readwrite("beta1", &NetworkConfig::beta1)
        .def_readwrite("beta2", &NetworkConfig::beta2)
        .def_readwrite("epsilon", &NetworkConfig::epsilon)
        .def_readwrite("l2_reg", &NetworkConfig::l2_reg)
        .def_readwrite("inference_fp16_features", &NetworkConfig::inference_fp16_features,
                       "mixed precision: render and update the occupancy grid on fp16 features, inputs rounded "
                       "to fp16 once (the reference's fp16 network); False: the training forward's split inputs")
        .def_property(
            "appearance_embedding_dims", [](const NetworkConfig& c) { return c.appearance_dims; },
            [](NetworkConfig& c, int v) {
                if (v != 0 && v != TN_APPEARANCE_DIMS)
                    throw std::invalid_argument("appearance_embedding_dims is 0 or " +
                                                std::to_string(TN_APPEARANCE_DIMS));
                c.appearance_dims = v;
            },
            "0: no appearance embedding; 16: one per training image (the reference's network)");

    py::class_<TrainingConfig>(m, "TrainingConfig")
        .def(py::init<>())
        .def_readwrite("network", &TrainingConfig::network)
        .def_readwrite("rng_seed", &TrainingConfig::rng_seed)
        .def_readwrite("distortion_lambda", &TrainingConfig::distortion_lambda)
        .def_readwrite("occupancy_threshold", &TrainingConfig::occupancy_threshold)
        .def_readwrite("ray_capacity", &TrainingConfig::ray_capacity);

    py::class_<NeRFTrainingController::TrainingMetrics>(m, "TrainingMetrics")
        .def_readonly("step", &NeRFTrainingController::TrainingMetrics::step)
        .def_readonly("loss", &NeRFTrainingController::TrainingMetrics::loss)
        .def_readonly("n_rays", &NeRFTrainingController::TrainingMetrics::n_rays)
        .def_readonly("n_samples", &NeRFTrainingController::TrainingMetrics::n_samples);

    py::class_<NeRFTrainingController::OccupancyGridMetrics>(m, "OccupancyGridMetrics")
        .def_readonly("n_occupied", &NeRFTrainingController::OccupancyGridMetrics::n_occupied)
        .def_readonly("n_total", &NeRFTrainingController::OccupancyGridMetrics::n_total);

    py::class_<NeRFTrainingController, std::shared_ptr<NeRFTrainingController>>(m, "Trainer")
        .def(py::init<NeRFProxy*, TrainingConfig>(), py::arg("nerf"), py::arg("config") = TrainingConfig(),
             py::keep_alive<1, 2>())  // the trainer keeps its NeRF alive
        .def("setup_data", &NeRFTrainingController::setup_data, py::arg("batch_size"),
             "Call this once before starting training.")
        .def("teardown", &NeRFTrainingController::teardown, "Call this before destroying the controller.")
        .def("reset_training", &NeRFTrainingController::reset_training, "Call this to clear the training data.")
        .def(
            "load_images",
            [](NeRFTrainingController& tc, py::object cb) {
                if (py::isinstance<py::function>(cb)) {
                    auto f = std::make_shared<py::function>(cb.cast<py::function>());
                    py::gil_scoped_release release;
                    tc.load_images([f](int n_loaded, int n_total) {
                        py::gil_scoped_acquire acquire;
                        (*f)(n_loaded, n_total);
                    });
                } else {
                    py::gil_scoped_release release;
                    tc.load_images();
                }
            },
            py::arg("on_image_loaded") = py::none())
        .def("update_occupancy_grid", &NeRFTrainingController::update_occupancy_grid, py::arg("training_step"))
        .def("train_step", &NeRFTrainingController::train_step)
        .def("train_steps", &NeRFTrainingController::train_steps, py::arg("n_steps"), py::arg("read_metrics") = true,
             py::arg("graph") = false, py::call_guard<py::gil_scoped_release>())
        .def("set_occupancy_from_density",
             [](NeRFTrainingController& tc, py::array_t<float, py::array::c_style | py::array::forcecast> d,
                float thr) { tc.set_occupancy_from_density(d.data(), thr); },
             py::arg("density"), py::arg("threshold"))
        .def("get_params",
             [](const NeRFTrainingController& tc) {
                 auto v = tc.get_params();
                 py::array_t<float> a(v.size());
                 std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(float));
                 return a;
             })
        .def("set_params",
             [](NeRFTrainingController& tc, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 tc.set_params(std::vector<float>(a.data(), a.data() + a.size()));
             })
        .def("get_cuda_memory_allocated", &NeRFTrainingController::get_cuda_memory_allocated)
        .def("get_occupancy_bitfield",
             [](NeRFTrainingController& tc) {
                 tc.synchronize();
                 const auto& og = tc.proxy->nerfs[0].occupancy_grid;
                 py::array_t<uint8_t> a(og.bitfield.size());
                 TN_HIP(hipMemcpy(a.mutable_data(), og.bitfield.data(), og.bitfield.size(), hipMemcpyDeviceToHost));
                 return a;
             })
        .def("get_occupancy_density",
             [](NeRFTrainingController& tc) {
                 tc.synchronize();
                 const auto& og = tc.proxy->nerfs[0].occupancy_grid;
                 py::array_t<float> a(og.density.size());
                 TN_HIP(hipMemcpy(a.mutable_data(), og.density.data(), og.density.bytes(), hipMemcpyDeviceToHost));
                 return a;
             })
        .def("set_timing", &NeRFTrainingController::set_timing, py::arg("enabled"))
        .def("phase_timings", &NeRFTrainingController::phase_timings)
        .def("reset_timings", &NeRFTrainingController::reset_timings)
        .def("samples_trained", &NeRFTrainingController::samples_trained)
        .def("synchronize", &NeRFTrainingController::synchronize, py::call_guard<py::gil_scoped_release>())
        // test support: the batch the last step trained on, the network
        // workspace of that step and the optimiser state (host copies)
        // which = 0: the batch the last step trained on; 1: the batch
        // generated ahead for the next step
        .def("_get_batch",
             [](NeRFTrainingController& tc, int which) {
                 tc.synchronize();
                 Trainer::Context* c = tc.context();
                 if (!c) throw std::runtime_error("call setup_data() first");
                 const TrainingWorkspace& w = c->batches[which ? c->cur : c->last];
                 py::dict d;
                 d["info"] = to_numpy(w.info);
                 d["random"] = to_numpy(w.random_float);
                 d["pix_rgba"] = to_numpy(w.pix_rgba);
                 d["ori"] = to_numpy(w.ray_origin);
                 d["dir"] = to_numpy(w.ray_dir);
                 d["t"] = to_numpy(w.ray_t);
                 d["tmax"] = to_numpy(w.ray_t_max);
                 d["img_id"] = to_numpy(w.ray_img_id);
                 d["alive"] = to_numpy(w.ray_alive);
                 d["n_steps"] = to_numpy(w.ray_step);
                 d["offset"] = to_numpy(w.ray_offset);
                 d["sample_t"] = to_numpy(w.sample_t);
                 d["sample_pos"] = to_numpy(w.sample_pos);
                 d["sample_dir"] = to_numpy(w.sample_dir);
                 d["sample_dt"] = to_numpy(w.sample_dt);
                 d["sample_m_norm"] = to_numpy(w.sample_m_norm);
                 d["sample_dt_norm"] = to_numpy(w.sample_dt_norm);
                 d["n_rays_next"] = to_numpy(c->n_rays_next);
                 d["ray_capacity"] = w.ray_capacity;
                 d["batch_size"] = w.batch_size;
                 return d;
             },
             py::arg("which") = 0)
        .def("_get_network_outputs",
             [](NeRFTrainingController& tc) {
                 tc.synchronize();
                 const NerfNetwork& n = tc.proxy->nerfs[0].network;
                 const int p = n.config.precision;
                 py::dict d;
                 d["density"] = typed_to_float(n.workspace.density, p);
                 d["rgb"] = typed_to_float(n.workspace.rgb, p);
                 d["dL_drgb"] = typed_to_float(n.workspace.dL_drgb, p);
                 d["dL_ds"] = to_numpy(n.workspace.dL_ds);
                 d["loss"] = to_numpy(n.workspace.loss_out);
                 return d;
             })
        .def("_graph_captures", [](NeRFTrainingController& tc) { return tc.context() ? tc.context()->n_captures : 0u; })
        .def(
            "autotune",
            [](NeRFTrainingController& tc, uint32_t rounds, uint32_t reps) {
                AutotuneResult r;
                {
                    py::gil_scoped_release release;
                    r = Autotuner::tune([&](uint32_t n) { tc.train_steps(n, false, true); },
                                        [&]() { return tc.samples_trained(); }, rounds, reps);
                }
                return autotune_dict(r);
            },
            py::arg("rounds_per_trial") = 16, py::arg("repetitions") = 3,
            "Pick the kernels' launch shapes by timing captured training steps (the steps train the NeRF).")
        .def("_get_optimizer_state",
             [](NeRFTrainingController& tc) {
                 tc.synchronize();
                 const NetworkParamsWorkspace& p = tc.proxy->nerfs[0].params;
                 py::dict d;
                 d["m1"] = to_numpy(p.m1);
                 d["m2"] = to_numpy(p.m2);
                 d["steps"] = to_numpy(p.steps);
                 d["n_grid"] = p.n_grid;
                 return d;
             })
        .def("_set_optimizer_state",
             [](NeRFTrainingController& tc, py::array_t<float, py::array::c_style | py::array::forcecast> m1,
                py::array_t<float, py::array::c_style | py::array::forcecast> m2,
                py::array_t<uint32_t, py::array::c_style | py::array::forcecast> st) {
                 tc.synchronize();
                 NetworkParamsWorkspace& p = tc.proxy->nerfs[0].params;
                 from_numpy(p.m1, m1);
                 from_numpy(p.m2, m2);
                 from_numpy(p.steps, st);
             })
        .def_readwrite("alpha_selection_threshold", &NeRFTrainingController::alpha_selection_threshold)
        .def_readwrite("alpha_selection_probability", &NeRFTrainingController::alpha_selection_probability)
        .def_readwrite("min_step_size", &NeRFTrainingController::min_step_size)
        .def_readwrite("use_distortion_loss", &NeRFTrainingController::use_distortion_loss)
        .def_property(
            "distortion_lambda", [](const NeRFTrainingController& tc) { return tc.config.distortion_lambda; },
            [](NeRFTrainingController& tc, float v) { tc.config.distortion_lambda = v; })
        .def_readwrite("nerf", &NeRFTrainingController::proxy);

    // several NeRFs trained side by side, one captured launch per round
    py::class_<NeRFMultiTrainingController>(m, "MultiTrainer")
        .def(py::init([](const std::vector<std::shared_ptr<NeRFTrainingController>>& ts, bool one_launch) {
                 std::vector<NeRFTrainingController*> raw;
                 for (const auto& t : ts) raw.push_back(t.get());
                 return std::make_unique<NeRFMultiTrainingController>(raw, one_launch);
             }),
             py::arg("trainers"), py::arg("one_launch") = false, py::keep_alive<1, 2>())
        .def("train_steps", &NeRFMultiTrainingController::train_steps, py::arg("n_rounds"),
             "One training step of every NeRF per round, as replays of captured graphs (each NeRF's own, or "
             "with one_launch one graph per round).",
             py::call_guard<py::gil_scoped_release>())
        .def_readonly("one_launch", &NeRFMultiTrainingController::one_launch)
        .def("synchronize", &NeRFMultiTrainingController::synchronize, py::call_guard<py::gil_scoped_release>())
        .def("samples_trained", &NeRFMultiTrainingController::samples_trained)
        .def(
            "autotune",
            [](NeRFMultiTrainingController& mt, uint32_t rounds, uint32_t reps) {
                AutotuneResult r;
                {
                    py::gil_scoped_release release;
                    r = Autotuner::tune([&](uint32_t n) { mt.train_steps(n); }, [&]() { return mt.samples_trained(); },
                                        rounds, reps);
                }
                return autotune_dict(r);
            },
            py::arg("rounds_per_trial") = 16, py::arg("repetitions") = 3,
            "Pick the kernels' launch shapes by timing rounds of these NeRFs (the rounds train them).")
        .def_readonly("_graph_captures", &NeRFMultiTrainingController::n_captures);

    m.def("get_tuning", &Autotuner::current, "Current launch-shape knobs (name -> value).");
    m.def("set_tuning", &Autotuner::set, py::arg("name"), py::arg("value"), "Set one launch-shape knob.");

    register_renderer_bindings(m);
}

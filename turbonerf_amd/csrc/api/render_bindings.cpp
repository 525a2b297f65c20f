// Rendering part of the PyTurboNeRF module (reference python_bindings.cu:292-419).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "renderer.hpp"

namespace py = pybind11;

namespace turbo {

void register_renderer_bindings(py::module_& m) {
    py::class_<SpatialEffect, std::shared_ptr<SpatialEffect>>(m, "SpatialEffect");
    py::class_<RepeaterEffect, SpatialEffect, std::shared_ptr<RepeaterEffect>>(m, "RepeaterEffect")
        .def(py::init<BoundingBox, BoundingBox, Transform4f>(), py::arg("source_bbox"), py::arg("extend_bbox"),
             py::arg("transform") = Transform4f::Identity())
        .def_readonly("source_bbox", &RepeaterEffect::source_bbox)
        .def_readonly("extend_bbox", &RepeaterEffect::extend_bbox);

    py::class_<NeRFRenderable>(m, "Renderable")
        .def(py::init([](NeRFProxy* nerf, std::vector<std::shared_ptr<SpatialEffect>> fx, py::object bbox) {
                 if (bbox.is_none()) return NeRFRenderable(nerf, std::move(fx));
                 return NeRFRenderable(nerf, bbox.cast<BoundingBox>(), std::move(fx));
             }),
             py::arg("nerf"), py::arg("spatial_effects") = std::vector<std::shared_ptr<SpatialEffect>>(),
             py::arg("bounding_box") = py::none(), py::keep_alive<1, 2>())
        .def_readonly("nerf", &NeRFRenderable::proxy)
        .def_readonly("spatial_effects", &NeRFRenderable::spatial_effects)
        .def_readwrite("bounding_box", &NeRFRenderable::bounding_box);

    py::enum_<RenderPattern>(m, "RenderPattern")
        .value("LinearChunks", RenderPattern::LinearChunks)
        .value("HexagonalGrid", RenderPattern::HexagonalGrid)
        .value("RectangularGrid", RenderPattern::RectangularGrid);

    py::enum_<RenderFlags>(m, "RenderFlags", py::arithmetic())
        .value("Preview", RenderFlags::Preview)
        .value("Final", RenderFlags::Final)
        .def(py::self | py::self)
        .def(py::self & py::self)
        .def(py::self |= py::self)
        .def(py::self &= py::self);

    py::class_<RenderProperties>(m, "RenderProperties")
        .def(py::init<>())
        .def_readwrite("show_near_planes", &RenderProperties::show_near_planes)
        .def_readwrite("show_far_planes", &RenderProperties::show_far_planes);

    py::class_<RenderModifiers>(m, "RenderModifiers")
        .def(py::init<>())
        .def_readwrite("properties", &RenderModifiers::properties);

    py::class_<RenderTarget>(m, "RenderTarget")
        .def("set_size", [](RenderTarget& rt, int w, int h) { rt.set_size(w, h); }, py::arg("width"), py::arg("height"))
        .def_readonly("width", &RenderTarget::width)
        .def_readonly("height", &RenderTarget::height)
        .def("free", [](RenderTarget& rt) { rt.free(); })
        .def("save_image", [](RenderTarget& rt, const std::string& path) { rt.save_image(path); },
             py::arg("file_path"));

    py::class_<CPURenderBuffer, RenderTarget>(m, "CPURenderBuffer")
        .def(py::init<>())
        .def("get_rgba", [](py::object self) {
            auto& rb = self.cast<CPURenderBuffer&>();
            const py::ssize_t w = rb.width, h = rb.height;
            // a view into the buffer, as the reference returns (python_bindings.cu:334-349)
            return py::array_t<float>({h, w, (py::ssize_t)4},
                                      {(py::ssize_t)(sizeof(float) * 4 * w), (py::ssize_t)(sizeof(float) * 4),
                                       (py::ssize_t)sizeof(float)},
                                      rb.get_rgba(), self);
        });

    py::class_<CUDARenderBuffer, RenderTarget>(m, "CUDARenderBuffer")
        .def(py::init<>())
        .def_property_readonly("data_ptr", [](CUDARenderBuffer& rb) { return reinterpret_cast<uintptr_t>(rb.data()); },
                               "device address of the interleaved float RGBA image")
        .def("get_rgba", [](CUDARenderBuffer& rb) {
            // a host copy (the image itself stays on the device)
            py::array_t<float> a({(py::ssize_t)rb.height, (py::ssize_t)rb.width, (py::ssize_t)4});
            if (rb.rgba.size())
                TN_HIP(hipMemcpy(a.mutable_data(), rb.rgba.data(), rb.rgba.bytes(), hipMemcpyDeviceToHost));
            return a;
        });

    py::class_<RenderRequest, std::shared_ptr<RenderRequest>>(m, "RenderRequest")
        .def(py::init([](const Camera& cam, std::vector<NeRFRenderable> nerfs, RenderTarget* output,
                         const RenderModifiers& mods, RenderFlags flags, OnCompleteCallback oc, OnProgressCallback op,
                         OnCancelCallback ocn) {
                 return std::make_shared<RenderRequest>(cam, std::move(nerfs), output, mods, flags, oc, op, ocn);
             }),
             py::arg("camera"), py::arg("nerfs"), py::arg("output"), py::arg("modifiers") = RenderModifiers(),
             py::arg("flags") = RenderFlags::Final, py::arg("on_complete") = nullptr, py::arg("on_progress") = nullptr,
             py::arg("on_cancel") = nullptr, py::keep_alive<1, 4>())
        .def("cancel", &RenderRequest::cancel);

    py::class_<NeRFRenderingController, std::shared_ptr<NeRFRenderingController>>(m, "Renderer")
        .def(py::init<RenderPattern, uint32_t>(), py::arg("pattern") = RenderPattern::LinearChunks,
             py::arg("batch_size") = 0)
        .def("submit", &NeRFRenderingController::submit, py::arg("request"))
        .def("cancel", &NeRFRenderingController::cancel)
        .def_readwrite("min_step_size", &NeRFRenderingController::min_step_size)
        .def_readwrite("max_samples_per_ray", &NeRFRenderingController::max_samples_per_ray)
        .def_readonly("last_sample_count", &NeRFRenderingController::last_sample_count)
        .def_readwrite("_force_shared_steps", &NeRFRenderingController::force_shared_steps);
}

}  // namespace turbo

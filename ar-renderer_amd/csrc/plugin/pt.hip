// Wavefront path tracer (placeholder until the kernels land).
#include "pt_host.h"

namespace pgpu {
struct PTScene {};
Device::Device() = default;
Device::~Device() {
    pt.reset();
    if (stream) (void)hipStreamDestroy(stream);
}
void ReleasePT(Device &d) { d.pt.reset(); }
void RenderExImpl(Device &, const prgpu::SceneDescEx &, const prgpu::RenderParamsEx &, double *, double *) {
    throw std::runtime_error("render_ex not implemented");
}
void SampleRadianceImpl(Device &, const prgpu::SceneDescEx &, const prgpu::RenderParamsEx &, prgpu::SampleRecordEx *) {
    throw std::runtime_error("sample_radiance not implemented");
}
void RenderV4Impl(Device &, const prgpu::SceneDesc &, const prgpu::RenderParams &, float *) {
    throw std::runtime_error("v4 render not implemented");
}
}  // namespace pgpu

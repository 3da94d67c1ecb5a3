// Host-side entry points of the wavefront path tracer and the geometry helpers it shares with
// the ray-query path.
#pragma once

#include "device.h"
#include "trace_dev.h"

namespace pgpu {

uint64_t GeometryKey(const prgpu::Primitive *prims, int count);
void EnsureGeometry(Device &d, const prgpu::Primitive *prims, int count);
TraceCtx MakeTraceCtx(Device &d);
int TraceGrid(const Device &d, int64_t n);
void CheckTraceFlags(Device &d);

void ReleasePT(Device &d);
// sheenTable: the sheen albedo table (16 x 32) to use instead of the built-in one, or null.
void RenderExImpl(Device &d, const prgpu::SceneDescEx &sd, const prgpu::RenderParamsEx &rp, double *sum,
                  double *sumsq, const float *sheenTable = nullptr);
void SampleRadianceImpl(Device &d, const prgpu::SceneDescEx &sd, const prgpu::RenderParamsEx &rp,
                        prgpu::SampleRecordEx *out);
void RenderV4Impl(Device &d, const prgpu::SceneDesc &sd, const prgpu::RenderParams &rp, float *outXYZ);

}  // namespace pgpu

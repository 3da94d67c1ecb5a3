// Oracle C API -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference CPU path tracer (src/integrators/path.cpp and the code
// under it), driven by the same flat scene description as the GPU plugin. Only tests/, the
// cpu_baseline leg of bench.py and __graft_entry__.smoke() load liboracle.so, and only as the
// checker. Nothing in the product links it.
#pragma once

#include "../include/prender_gpu.h"

extern "C" {
void *or_scene_create(const prgpu::SceneDescEx *d);
void or_scene_destroy(void *h);
int or_scene_info(void *h, int64_t *out);  // [bvh nodes, light-bvh nodes, infinite lights, film margin]
void or_world_sphere(void *h, float *out);
int or_bvh_dump(void *h, float *nodes, int32_t *order);
void or_trace(void *h, const prgpu::RayEx *rays, int64_t n, int anyHit, prgpu::HitEx *hits, int threads);
void or_trace_brute(void *h, const prgpu::RayEx *rays, int64_t n, prgpu::HitEx *hits, int32_t *ties, int threads);
int or_light_bvh_dump(void *h, float *f, int32_t *iv);
int or_light_sample(void *h, int useBvh, const float *p, const float *n, float u, float *pmf);
float or_light_pmf(void *h, int useBvh, const float *p, const float *n, int light);
void or_camera_ray(void *h, float fx, float fy, float lx, float ly, float *out);
void or_sample_radiance(void *h, const prgpu::RenderParamsEx *params, prgpu::SampleRecordEx *out, int threads);
double or_render(void *h, const prgpu::RenderParamsEx *params, double *sum, double *sumsq, int threads);

void or_rng_stream(uint64_t pixel, uint64_t sampleIndex, uint64_t seed, int n, float *out);
void or_sample_visible(float u, float *l, float *pdf);
void or_cie(float l, float *xyz);
float or_cie_y_integral();
void or_to_xyz(const float *L, const float *l, const float *pdf, float *xyz);
float or_spectrum_eval(const prgpu::SpectrumEx *s, const float *knots, int nknots, float l);
float or_spectrum_max(const prgpu::SpectrumEx *s, const float *knots, int nknots);
float or_spectrum_y(const prgpu::SpectrumEx *s, const float *knots, int nknots);
void or_bxdf(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, float *out);
// Plain, coated (layered) and sheen closures with transport mode (0 radiance, 1 importance) and
// sample flags (1 reflection, 2 transmission, 3 all); descriptor layout in pt_oracle.cpp.
void or_closure(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, int mode, int rt,
                float *out);
float or_sheen_albedo(float cosTheta, float alpha);
void or_sheen_table(float *out);  // 16 alphas x 32 cosines
float or_fr_dielectric(float c, float eta);
float or_fr_complex(float c, float eta, float k);
float or_hg(float c, float g);
void or_sample_hg(const float *wo, float g, const float *u, float *wi, float *pdf);
void or_filter(int type, float radius, const float *x, int n, float *out, float *norm);
void or_light_sample_li(void *h, int light, const float *p, const float *n, const float *u, const float *lam,
                        const float *lpdf, float *out);
float or_light_pdf_li(void *h, int light, const float *p, const float *wi);
float or_light_power(void *h, int light);
void or_medium_tr(void *h, int medium, const float *o, const float *d, float tMax, const float *lam, const float *lpdf,
                  float *out);
void or_medium_sample(void *h, int medium, const float *o, const float *d, float tMax, float u, const float *lam,
                      const float *lpdf, float *out);

// Path-guiding model in FP64 (guide_oracle.cpp).
int or_guide_param_count();
void or_guide_forward(const float *params, const float *recs9, int n, float *out41);
double or_guide_mix_pdf(const float *out41, const float *w, double *sel);
void or_guide_mix_sample(const float *out41, float uLobe, float u1, float u2, float *w);
double or_guide_loss_grad(const float *params, const float *recs16, int n, float *grad);
}

// Screened Poisson reconstruction behind ar_poisson_solve
// (agent-ray src/integrators/gradient_domain_pt.rs:449-630), solved by
// conjugate gradient on the GPU instead of the reference's single-threaded
// red-black SOR.
//
// System (per colour channel, Neumann boundary, gradient_domain_pt.rs:503-595):
//   (alpha + n_neigh) I(x,y) - sum_neigh I = alpha P(x,y) + sign * div G
//   div G = dx(x,y) - dx(x-1,y)[x>0] + dy(x,y) - dy(x,y-1)[y>0]
// sign = +1 is the reference's right-hand side; sign = -1 the textbook
// normal equations of alpha|I-P|^2 + |DI-G|^2.  Warm start I = P, output
// clamped at 0 (gradient_domain_pt.rs:616-627).
//
// Vectors are fp32, one plane per channel (12 B/pixel/vector).  Dot products
// and the convergence test accumulate in fp64 through a two-level,
// fixed-order reduction (per-block partials, then one block), so the result
// is deterministic.  Scalars (alpha_k, beta_k, residuals) stay on the device,
// so the iteration loop runs without host synchronisation.
//
// Per iteration, two bandwidth-bound kernels:
//   A: p = r + beta p_prev (stencil of 5), q = A p, partial p.q
//      reads r, p_prev; writes p, q             48 B/pixel
//   B: x += a p, r -= a q, partial r.r
//      reads x, p, r, q; writes x, r            72 B/pixel
#include "ar_render.h"

namespace ar {

constexpr int kPoisBlock = 256;
constexpr int kPoisMaxBlocks = 1024;

struct PoisScalars {
  double rr[3];      // current r.r per channel
  double pq[3];
  double alpha[3], beta[3];
  double bb[3];      // |b|^2 (convergence reference)
  double rr0[3];
  int done;          // set when every channel has converged
  int iters_run;
};

struct PoisBufs {
  const float* P;
  const float* gx;
  const float* gy;
  float* x;
  float* r;
  float* p[2];
  float* q;
  float* b;
  double* partial;   // [3][kPoisMaxBlocks]
  PoisScalars* sc;
  uint32_t w, h, n;
  float alpha;
  int sign;
  double tol2;
};

__device__ __forceinline__ float stencil_apply(const float* v, uint32_t X, uint32_t Y, uint32_t w, uint32_t h,
                                               float alpha) {
  // (alpha + n_neigh) v - sum_neigh v, neighbours summed in the reference
  // order (left, right, up, down; gradient_domain_pt.rs:565-592).
  size_t i = (size_t)Y * w + X;
  float sum = 0.0f;
  uint32_t nn = 0;
  if (X > 0) { sum += v[i - 1]; ++nn; }
  if (X + 1 < w) { sum += v[i + 1]; ++nn; }
  if (Y > 0) { sum += v[i - w]; ++nn; }
  if (Y + 1 < h) { sum += v[i + w]; ++nn; }
  return (alpha + (float)nn) * v[i] - sum;
}

__device__ __forceinline__ void block_reduce3(double a0, double a1, double a2, double* out0, int block_slot) {
  __shared__ double sh[3][kPoisBlock / 64];
  double v[3] = {a0, a1, a2};
  for (int c = 0; c < 3; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) sh[c][wave] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < kPoisBlock / 64; ++k) s += sh[c][k];
      out0[c * kPoisMaxBlocks + block_slot] = s;
    }
  }
}

// b = alpha P + sign div G; x = P; r = b - A x; p_prev = 0; partial r.r, b.b
__global__ __launch_bounds__(kPoisBlock) void k_pois_init(PoisBufs B) {
  double rr[3] = {0, 0, 0}, bb[3] = {0, 0, 0};
  for (uint32_t c = 0; c < 3; ++c) {
    const float* P = B.P + (size_t)c * B.n;
    const float* gx = B.gx + (size_t)c * B.n;
    const float* gy = B.gy + (size_t)c * B.n;
    for (uint32_t i = blockIdx.x * kPoisBlock + threadIdx.x; i < B.n; i += gridDim.x * kPoisBlock) {
      uint32_t X = i % B.w, Y = i / B.w;
      float div = gx[i];
      if (X > 0) div -= gx[i - 1];
      div += gy[i];
      if (Y > 0) div -= gy[i - B.w];
      float bv = B.alpha * P[i] + (B.sign >= 0 ? div : -div);
      B.b[(size_t)c * B.n + i] = bv;
      B.x[(size_t)c * B.n + i] = P[i];
      B.p[0][(size_t)c * B.n + i] = 0.0f;
      bb[c] += (double)bv * (double)bv;
    }
  }
  // second pass needs x complete only at neighbours owned by other threads:
  // residual is formed in k_pois_residual after a grid-wide boundary.
  block_reduce3(bb[0], bb[1], bb[2], B.partial, blockIdx.x);
  (void)rr;
}

__global__ __launch_bounds__(kPoisBlock) void k_pois_residual(PoisBufs B) {
  double rr[3] = {0, 0, 0};
  for (uint32_t c = 0; c < 3; ++c) {
    const float* x = B.x + (size_t)c * B.n;
    for (uint32_t i = blockIdx.x * kPoisBlock + threadIdx.x; i < B.n; i += gridDim.x * kPoisBlock) {
      uint32_t X = i % B.w, Y = i / B.w;
      float r = B.b[(size_t)c * B.n + i] - stencil_apply(x, X, Y, B.w, B.h, B.alpha);
      B.r[(size_t)c * B.n + i] = r;
      rr[c] += (double)r * (double)r;
    }
  }
  block_reduce3(rr[0], rr[1], rr[2], B.partial, blockIdx.x);
}

// One block: fixed-order sum of per-block partials.  mode 0: bb and rr0
// setup (called twice), mode 1: alpha = rr / pq, mode 2: beta = rr_new / rr.
// The partials are summed by kPoisBlock threads in a fixed order (thread j
// takes partials j, j + kPoisBlock, ...; then a fixed shuffle tree per wave
// and the waves in order), so the scalars are deterministic.
__global__ __launch_bounds__(kPoisBlock) void k_pois_scalar(PoisBufs B, int nblocks, int mode) {
  __shared__ double sh[3][kPoisBlock / 64];
  __shared__ double tot[3];
  double v[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < nblocks; k += kPoisBlock)
    for (int c = 0; c < 3; ++c) v[c] += B.partial[c * kPoisMaxBlocks + k];
  for (int c = 0; c < 3; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) sh[c][wave] = v[c];
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int k = 0; k < kPoisBlock / 64; ++k) t += sh[threadIdx.x][k];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  PoisScalars* S = B.sc;
  if (mode == 10) {  // after init: |b|^2
    for (int c = 0; c < 3; ++c) S->bb[c] = tot[c];
    return;
  }
  if (mode == 11) {  // after first residual: r.r, beta = 0, not done
    for (int c = 0; c < 3; ++c) {
      S->rr[c] = tot[c];
      S->rr0[c] = tot[c];
      S->beta[c] = 0.0;
    }
    S->done = 0;
    S->iters_run = 0;
    bool all = true;
    for (int c = 0; c < 3; ++c)
      if (!(S->rr[c] <= B.tol2 * S->bb[c])) all = false;
    if (all) S->done = 1;
    return;
  }
  if (S->done) return;
  if (mode == 1) {
    for (int c = 0; c < 3; ++c) {
      S->pq[c] = tot[c];
      S->alpha[c] = tot[c] != 0.0 ? S->rr[c] / tot[c] : 0.0;
    }
    return;
  }
  // mode 2: new r.r
  bool all = true;
  for (int c = 0; c < 3; ++c) {
    double rn = tot[c];
    S->beta[c] = S->rr[c] != 0.0 ? rn / S->rr[c] : 0.0;
    S->rr[c] = rn;
    if (!(rn <= B.tol2 * S->bb[c])) all = false;
  }
  S->iters_run += 1;
  if (all) S->done = 1;
}

// A: p = r + beta p_prev; q = A p; partial p.q.  The stencil needs p at the
// neighbours, which this kernel computes itself from r and p_prev.
__global__ __launch_bounds__(kPoisBlock) void k_pois_A(PoisBufs B, int pin) {
  if (B.sc->done) return;
  double pq[3] = {0, 0, 0};
  const uint32_t w = B.w, h = B.h;
  for (uint32_t c = 0; c < 3; ++c) {
    const float beta = (float)B.sc->beta[c];
    const float* r = B.r + (size_t)c * B.n;
    const float* pp = B.p[pin] + (size_t)c * B.n;
    float* pn = B.p[pin ^ 1] + (size_t)c * B.n;
    float* q = B.q + (size_t)c * B.n;
    for (uint32_t i = blockIdx.x * kPoisBlock + threadIdx.x; i < B.n; i += gridDim.x * kPoisBlock) {
      uint32_t X = i % w, Y = i / w;
      float pc = r[i] + beta * pp[i];
      float sum = 0.0f;
      uint32_t nn = 0;
      if (X > 0) { sum += r[i - 1] + beta * pp[i - 1]; ++nn; }
      if (X + 1 < w) { sum += r[i + 1] + beta * pp[i + 1]; ++nn; }
      if (Y > 0) { sum += r[i - w] + beta * pp[i - w]; ++nn; }
      if (Y + 1 < h) { sum += r[i + w] + beta * pp[i + w]; ++nn; }
      float qv = (B.alpha + (float)nn) * pc - sum;
      pn[i] = pc;
      q[i] = qv;
      pq[c] += (double)pc * (double)qv;
    }
  }
  block_reduce3(pq[0], pq[1], pq[2], B.partial, blockIdx.x);
}

// B: x += alpha p; r -= alpha q; partial r.r
__global__ __launch_bounds__(kPoisBlock) void k_pois_B(PoisBufs B, int pcur) {
  if (B.sc->done) return;
  double rr[3] = {0, 0, 0};
  for (uint32_t c = 0; c < 3; ++c) {
    const float a = (float)B.sc->alpha[c];
    float* x = B.x + (size_t)c * B.n;
    float* r = B.r + (size_t)c * B.n;
    const float* p = B.p[pcur] + (size_t)c * B.n;
    const float* q = B.q + (size_t)c * B.n;
    for (uint32_t i = blockIdx.x * kPoisBlock + threadIdx.x; i < B.n; i += gridDim.x * kPoisBlock) {
      x[i] = x[i] + a * p[i];
      float rv = r[i] - a * q[i];
      r[i] = rv;
      rr[c] += (double)rv * (double)rv;
    }
  }
  block_reduce3(rr[0], rr[1], rr[2], B.partial, blockIdx.x);
}

// out (interleaved RGB) = max(x, 0)
__global__ void k_pois_out(PoisBufs B, float* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = fmaxf(B.x[(size_t)c * B.n + i], 0.0f);
}

// interleaved RGB -> planes
__global__ void k_deinterleave(const float* in, float* out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = in[3 * (size_t)i + c];
}

size_t poisson_scratch_bytes(uint32_t w, uint32_t h) {
  size_t plane = (size_t)w * h * 3 * sizeof(float);
  size_t pl_al = (plane + 255) / 256 * 256;
  return pl_al * 9 + 3 * kPoisMaxBlocks * sizeof(double) + 256;
}

// Solve with interleaved RGB device inputs; scratch is caller-provided.
ar_status poisson_solve_device(const float* P_rgb, const float* gx_rgb, const float* gy_rgb, uint32_t w, uint32_t h,
                               float alpha, uint32_t iters, int32_t sign, float* out_rgb, double* final_residual,
                               DevBuf& scratch, hipStream_t st) {
  uint32_t n = w * h;
  size_t plane = (size_t)n * 3 * sizeof(float);
  size_t pl_al = (plane + 255) / 256 * 256;
  size_t need = poisson_scratch_bytes(w, h);
  if (scratch.bytes < need) {
    AR_HIP(hipStreamSynchronize(st));
    AR_HIP(scratch.alloc(need));
  }
  char* b = (char*)scratch.p;
  float* planes[9];
  for (int i = 0; i < 9; ++i) planes[i] = (float*)(b + pl_al * i);
  PoisBufs B;
  B.P = planes[0];
  B.gx = planes[1];
  B.gy = planes[2];
  B.x = planes[3];
  B.r = planes[4];
  B.p[0] = planes[5];
  B.p[1] = planes[6];
  B.q = planes[7];
  B.b = planes[8];
  B.partial = (double*)(b + pl_al * 9);
  B.sc = (PoisScalars*)(b + pl_al * 9 + 3 * kPoisMaxBlocks * sizeof(double));
  B.w = w;
  B.h = h;
  B.n = n;
  B.alpha = alpha;
  B.sign = sign;
  B.tol2 = 1e-24;  // run the requested iterations unless the residual is exhausted
  int grid_e = (int)std::min<uint32_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_deinterleave, dim3(grid_e), dim3(256), 0, st, P_rgb, planes[0], n);
  hipLaunchKernelGGL(k_deinterleave, dim3(grid_e), dim3(256), 0, st, gx_rgb, planes[1], n);
  hipLaunchKernelGGL(k_deinterleave, dim3(grid_e), dim3(256), 0, st, gy_rgb, planes[2], n);
  int nb = (int)std::min<uint32_t>((n + kPoisBlock - 1) / kPoisBlock, kPoisMaxBlocks);
  hipLaunchKernelGGL(k_pois_init, dim3(nb), dim3(kPoisBlock), 0, st, B);
  hipLaunchKernelGGL(k_pois_scalar, dim3(1), dim3(kPoisBlock), 0, st, B, nb, 10);
  hipLaunchKernelGGL(k_pois_residual, dim3(nb), dim3(kPoisBlock), 0, st, B);
  hipLaunchKernelGGL(k_pois_scalar, dim3(1), dim3(kPoisBlock), 0, st, B, nb, 11);
  int pin = 0;
  for (uint32_t it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_pois_A, dim3(nb), dim3(kPoisBlock), 0, st, B, pin);
    hipLaunchKernelGGL(k_pois_scalar, dim3(1), dim3(kPoisBlock), 0, st, B, nb, 1);
    hipLaunchKernelGGL(k_pois_B, dim3(nb), dim3(kPoisBlock), 0, st, B, pin ^ 1);
    hipLaunchKernelGGL(k_pois_scalar, dim3(1), dim3(kPoisBlock), 0, st, B, nb, 2);
    pin ^= 1;
  }
  hipLaunchKernelGGL(k_pois_out, dim3(grid_e), dim3(256), 0, st, B, out_rgb);
  AR_HIP(hipGetLastError());
  if (final_residual) {
    PoisScalars hs;
    AR_HIP(hipMemcpyAsync(&hs, B.sc, sizeof(hs), hipMemcpyDeviceToHost, st));
    AR_HIP(hipStreamSynchronize(st));
    double num = 0, den = 0;
    for (int c = 0; c < 3; ++c) {
      num += hs.rr[c];
      den += hs.bb[c];
    }
    *final_residual = den > 0 ? sqrt(num / den) : sqrt(num);
  }
  return AR_OK;
}

}  // namespace ar

using namespace ar;

extern "C" ar_status ar_poisson_solve(const float* primal, const float* dx, const float* dy, uint32_t width,
                                      uint32_t height, float alpha, uint32_t iterations, int32_t sign, float* out,
                                      double* final_residual, uint32_t flags, void* stream) {
  if (!primal || !dx || !dy || !out) return fail(AR_ERR_INVALID_ARG, "ar_poisson_solve: NULL argument");
  if (width == 0 || height == 0) return fail(AR_ERR_INVALID_ARG, "ar_poisson_solve: empty image");
  if (!(alpha > 0.0f)) return fail(AR_ERR_INVALID_ARG, "ar_poisson_solve: alpha must be > 0 (SPD system)");
  if ((uint64_t)width * height > (1ull << 28)) return fail(AR_ERR_INVALID_ARG, "ar_poisson_solve: image too large");
  hipStream_t st = (hipStream_t)stream;
  size_t bytes = (size_t)width * height * 3 * sizeof(float);
  DevBuf in, scratch;
  const float *dp = primal, *dgx = dx, *dgy = dy;
  float* dout = out;
  bool host = flags & AR_HOST_PTRS;
  if (host) {
    AR_HIP(in.alloc(bytes * 4));
    float* base = in.as<float>();
    size_t nf = bytes / sizeof(float);
    AR_HIP(hipMemcpyAsync(base, primal, bytes, hipMemcpyHostToDevice, st));
    AR_HIP(hipMemcpyAsync(base + nf, dx, bytes, hipMemcpyHostToDevice, st));
    AR_HIP(hipMemcpyAsync(base + 2 * nf, dy, bytes, hipMemcpyHostToDevice, st));
    dp = base;
    dgx = base + nf;
    dgy = base + 2 * nf;
    dout = base + 3 * nf;
  }
  ar_status e = poisson_solve_device(dp, dgx, dgy, width, height, alpha, iterations, sign, dout, final_residual,
                                     scratch, st);
  if (e != AR_OK) return e;
  if (host) AR_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, st));
  AR_HIP(hipStreamSynchronize(st));
  return AR_OK;
}

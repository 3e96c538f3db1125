// Multi-vector kernels of the FP64 GMRES refinement.  Every right-hand side
// (column t) runs its own Krylov process, so the Gram-Schmidt products are
// per-column dot products, not GEMMs:
//   mdot:  h[i][t] = sum_r conj(V_i[r][t]) w[r][t],   i < k
//   maxpy: w[r][t] += alpha sum_i V_i[r][t] h[i][t]
// with V_i = V + i sV an n x T row-major complex128 matrix.  Both passes are
// HBM-bound (16 k n T bytes read).  mdot sums fixed row chunks in a fixed
// order (partials, then one reduction kernel), so its result does not depend
// on scheduling: the same inputs give the same bits on every call.
#include "mg_internal.h"

namespace {

constexpr int kChunks = 128;   // row chunks of the dot-product pass
constexpr int kVec = 8;        // vectors accumulated per pass over w

__global__ __launch_bounds__(256) void k_mdot_part(const double2* __restrict__ V, int64_t sV, int64_t ldv, int k,
                                                   const double2* __restrict__ w, int64_t ldw, int64_t n, int64_t T,
                                                   int64_t rows, double2* __restrict__ part) {
  __shared__ double2 red[4][kVec][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t c = blockIdx.x, t = (int64_t)blockIdx.y * 64 + tx;
  const int64_t r0 = c * rows, r1 = (r0 + rows < n) ? r0 + rows : n;
  for (int i0 = 0; i0 < k; i0 += kVec) {
    double2 acc[kVec];
#pragma unroll
    for (int ii = 0; ii < kVec; ++ii) acc[ii] = make_double2(0.0, 0.0);
    if (t < T) {
      for (int64_t r = r0 + ty; r < r1; r += 4) {
        const double2 x = w[r * ldw + t];
#pragma unroll
        for (int ii = 0; ii < kVec; ++ii) {
          if (i0 + ii < k) {
            const double2 v = V[(int64_t)(i0 + ii) * sV + r * ldv + t];
            acc[ii].x += v.x * x.x + v.y * x.y;   // conj(v) x
            acc[ii].y += v.x * x.y - v.y * x.x;
          }
        }
      }
    }
#pragma unroll
    for (int ii = 0; ii < kVec; ++ii) red[ty][ii][tx] = acc[ii];
    __syncthreads();
    if (ty == 0 && t < T) {
#pragma unroll
      for (int ii = 0; ii < kVec; ++ii) {
        if (i0 + ii < k) {
          double2 s = red[0][ii][tx];
#pragma unroll
          for (int q = 1; q < 4; ++q) {
            s.x += red[q][ii][tx].x;
            s.y += red[q][ii][tx].y;
          }
          part[(c * k + i0 + ii) * T + t] = s;
        }
      }
    }
    __syncthreads();
  }
}

// The first Gram-Schmidt update fused with the second pass's dots: over row
// chunk blockIdx.x and 64 columns, w <- w - sum_i V_i h_i (the arithmetic of
// k_maxpy), then part[(c k + i) T + t] = the chunk's sum of conj(V_i) w (as
// k_mdot_part) with the updated w; V_i of a row is read once for both.
constexpr int kFuseMax = 16;   // vectors held in registers

__global__ __launch_bounds__(256) void k_maxpy_mdot(const double2* __restrict__ V, int64_t sV, int64_t ldv, int k,
                                                    const double2* __restrict__ h, int64_t ldh,
                                                    double2* __restrict__ w, int64_t ldw, int64_t n, int64_t T,
                                                    int64_t rows, double2* __restrict__ part) {
  __shared__ double2 red[4][kVec][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t c = blockIdx.x, t = (int64_t)blockIdx.y * 64 + tx;
  const int64_t r0 = c * rows, r1 = (r0 + rows < n) ? r0 + rows : n;
  double2 acc[kFuseMax], hv[kFuseMax];
#pragma unroll
  for (int i = 0; i < kFuseMax; ++i) {
    acc[i] = make_double2(0.0, 0.0);
    hv[i] = (i < k && t < T) ? h[(int64_t)i * ldh + t] : make_double2(0.0, 0.0);
  }
  if (t < T) {
    for (int64_t r = r0 + ty; r < r1; r += 4) {
      double2 v[kFuseMax];
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < kFuseMax; ++i) {
        if (i < k) {
          v[i] = V[(int64_t)i * sV + r * ldv + t];
          s.x += v[i].x * hv[i].x - v[i].y * hv[i].y;
          s.y += v[i].x * hv[i].y + v[i].y * hv[i].x;
        }
      }
      double2 x = w[r * ldw + t];
      x.x -= s.x;
      x.y -= s.y;
      w[r * ldw + t] = x;
#pragma unroll
      for (int i = 0; i < kFuseMax; ++i) {
        if (i < k) {
          acc[i].x += v[i].x * x.x + v[i].y * x.y;   // conj(v) x
          acc[i].y += v[i].x * x.y - v[i].y * x.x;
        }
      }
    }
  }
  for (int i0 = 0; i0 < k; i0 += kVec) {
#pragma unroll
    for (int ii = 0; ii < kVec; ++ii) {
      double2 a = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < kFuseMax; ++i)
        if (i == i0 + ii) a = acc[i];
      red[ty][ii][tx] = a;
    }
    __syncthreads();
    if (ty == 0 && t < T) {
#pragma unroll
      for (int ii = 0; ii < kVec; ++ii) {
        if (i0 + ii < k) {
          double2 o = red[0][ii][tx];
#pragma unroll
          for (int q = 1; q < 4; ++q) {
            o.x += red[q][ii][tx].x;
            o.y += red[q][ii][tx].y;
          }
          part[(c * k + i0 + ii) * T + t] = o;
        }
      }
    }
    __syncthreads();
  }
}

// out[i ldo + t] = sum over chunks of part: one workgroup per vector i and 64
// columns, its 4 waves summing interleaved chunks, combined in a fixed order.
__global__ __launch_bounds__(256) void k_mdot_reduce(const double2* __restrict__ part, int k, int64_t T, int nchunks,
                                                     double2* __restrict__ out, int64_t ldo) {
  __shared__ double2 red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * 64 + tx;
  double2 s = make_double2(0.0, 0.0);
  if (t < T) {
    for (int c = ty; c < nchunks; c += 4) {
      const double2 p = part[((int64_t)c * k + i) * T + t];
      s.x += p.x;
      s.y += p.y;
    }
  }
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && t < T) {
    double2 o = red[0][tx];
#pragma unroll
    for (int q = 1; q < 4; ++q) {
      o.x += red[q][tx].x;
      o.y += red[q][tx].y;
    }
    out[i * ldo + t] = o;
  }
}

// w += alpha sum_i V_i h_i (the arithmetic of k_maxpy) over row chunk
// blockIdx.x and 64 columns, with part[c T + t] = that chunk's sum of
// |w[r, t]|^2 after the update.
__global__ __launch_bounds__(256) void k_maxpy_norm(const double2* __restrict__ V, int64_t sV, int64_t ldv, int k,
                                                    const double2* __restrict__ h, int64_t ldh, double alpha,
                                                    double2* __restrict__ w, int64_t ldw, int64_t n, int64_t T,
                                                    int64_t rows, double* __restrict__ part) {
  __shared__ double red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t c = blockIdx.x, t = (int64_t)blockIdx.y * 64 + tx;
  const int64_t r0 = c * rows, r1 = (r0 + rows < n) ? r0 + rows : n;
  double acc = 0.0;
  if (t < T) {
    for (int64_t r = r0 + ty; r < r1; r += 4) {
      double2 s = make_double2(0.0, 0.0);
      for (int i = 0; i < k; ++i) {
        const double2 v = V[(int64_t)i * sV + r * ldv + t], q = h[(int64_t)i * ldh + t];
        s.x += v.x * q.x - v.y * q.y;
        s.y += v.x * q.y + v.y * q.x;
      }
      double2 o = w[r * ldw + t];
      o.x += alpha * s.x;
      o.y += alpha * s.y;
      w[r * ldw + t] = o;
      acc += o.x * o.x + o.y * o.y;
    }
  }
  red[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && t < T) part[c * T + t] = ((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx];
}

// part[c T + t] = sum over row chunk c of |X[r, t]|^2 (the chunks of k_maxpy_norm)
__global__ __launch_bounds__(256) void k_colnorm_part(const double2* __restrict__ X, int64_t ldx, int64_t n, int64_t T,
                                                      int64_t rows, double* __restrict__ part) {
  __shared__ double red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t c = blockIdx.x, t = (int64_t)blockIdx.y * 64 + tx;
  const int64_t r0 = c * rows, r1 = (r0 + rows < n) ? r0 + rows : n;
  double acc = 0.0;
  if (t < T)
    for (int64_t r = r0 + ty; r < r1; r += 4) {
      const double2 o = X[r * ldx + t];
      acc += o.x * o.x + o.y * o.y;
    }
  red[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && t < T) part[c * T + t] = ((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx];
}

// out[t] = sqrt(sum over chunks of part[c T + t]) (fixed order, as k_mdot_reduce)
__global__ __launch_bounds__(256) void k_norm_reduce(const double* __restrict__ part, int64_t T, int nchunks,
                                                     double* __restrict__ out) {
  __shared__ double red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t t = (int64_t)blockIdx.x * 64 + tx;
  double s = 0.0;
  if (t < T)
    for (int c = ty; c < nchunks; c += 4) s += part[(int64_t)c * T + t];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && t < T) out[t] = sqrt(((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx]);
}

__global__ __launch_bounds__(256) void k_maxpy(const double2* __restrict__ V, int64_t sV, int64_t ldv, int k,
                                               const double2* __restrict__ h, int64_t ldh, double alpha,
                                               double2* __restrict__ w, int64_t ldw, int64_t n, int64_t T) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * T) return;
  const int64_t r = e / T, t = e - r * T;
  double2 s = make_double2(0.0, 0.0);
  for (int i = 0; i < k; ++i) {
    const double2 v = V[(int64_t)i * sV + r * ldv + t], c = h[(int64_t)i * ldh + t];
    s.x += v.x * c.x - v.y * c.y;
    s.y += v.x * c.y + v.y * c.x;
  }
  double2 o = w[r * ldw + t];
  o.x += alpha * s.x;
  o.y += alpha * s.y;
  w[r * ldw + t] = o;
}

// Column j of the Hessenberg matrices of all right-hand sides, one thread
// per column t (H: (m+1) x m x T, entry (i, j) of column t at
// H[(i m + j) T + t]; cs, sn: m x T; g: (m+1) x T).  On entry H holds the
// Gram-Schmidt coefficients h_ij, i <= j; hn[t] = ||w|| becomes h_{j+1,j}.
// The unrotated column goes to H0, then the j earlier Givens rotations and a
// new one (which zeroes h_{j+1,j}) are applied and g is rotated with it.
// est[0] = max_t |g_{j+1,t}| / bnorm[t], the largest relative residual of the
// least-squares problem (one workgroup: a fixed reduction order).
__global__ __launch_bounds__(256) void k_givens(double2* __restrict__ H, double2* __restrict__ H0,
                                                double2* __restrict__ cs, double2* __restrict__ sn,
                                                double2* __restrict__ g, const double* __restrict__ hn,
                                                const double* __restrict__ bnorm, int m, int j, int64_t T,
                                                double* __restrict__ est) {
  __shared__ double red[256];
  double mx = 0.0;
  for (int64_t t = threadIdx.x; t < T; t += blockDim.x) {
    auto at = [&](int i) -> double2& { return H[((int64_t)i * m + j) * T + t]; };
    at(j + 1) = make_double2(hn[t], 0.0);
    for (int i = 0; i <= j + 1; ++i) H0[((int64_t)i * m + j) * T + t] = at(i);
    for (int i = 0; i < j; ++i) {
      const double2 c = cs[(int64_t)i * T + t], sv = sn[(int64_t)i * T + t];
      const double2 a = at(i), b = at(i + 1);
      // conj(c) a + conj(s) b  and  -s a + c b
      at(i) = make_double2(c.x * a.x + c.y * a.y + sv.x * b.x + sv.y * b.y,
                           c.x * a.y - c.y * a.x + sv.x * b.y - sv.y * b.x);
      at(i + 1) = make_double2(-(sv.x * a.x - sv.y * a.y) + (c.x * b.x - c.y * b.y),
                               -(sv.x * a.y + sv.y * a.x) + (c.x * b.y + c.y * b.x));
    }
    const double2 a = at(j), b = at(j + 1);
    double den = sqrt(a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y);
    if (den == 0.0) den = 1.0;
    const double2 c = make_double2(a.x / den, a.y / den), sv = make_double2(b.x / den, b.y / den);
    cs[(int64_t)j * T + t] = c;
    sn[(int64_t)j * T + t] = sv;
    at(j) = make_double2(den, 0.0);
    at(j + 1) = make_double2(0.0, 0.0);
    const double2 gj = g[(int64_t)j * T + t];
    const double2 g1 = make_double2(-(sv.x * gj.x - sv.y * gj.y), -(sv.x * gj.y + sv.y * gj.x));
    g[(int64_t)(j + 1) * T + t] = g1;
    g[(int64_t)j * T + t] = make_double2(c.x * gj.x + c.y * gj.y, c.x * gj.y - c.y * gj.x);
    const double e = sqrt(g1.x * g1.x + g1.y * g1.y) / bnorm[t];
    mx = (e > mx || e != e) ? e : mx;   // a NaN wins, so that it reaches the caller
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const double o = red[threadIdx.x + off];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) est[0] = red[0];
}

// y (k x T) from the Givens-rotated Hessenberg matrices: column t solves the
// upper-triangular R_t y_t = g_t (R_t = H[:k, :k] of column t) by back
// substitution, one thread per column.
__global__ __launch_bounds__(256) void k_hess_solve(const double2* __restrict__ H, const double2* __restrict__ g,
                                                    int m, int k, int64_t T, double2* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int i = k - 1; i >= 0; --i) {
    double2 s = g[(int64_t)i * T + t];
    for (int j = i + 1; j < k; ++j) {
      const double2 h = H[((int64_t)i * m + j) * T + t], yj = y[(int64_t)j * T + t];
      s.x -= h.x * yj.x - h.y * yj.y;
      s.y -= h.x * yj.y + h.y * yj.x;
    }
    const double2 d = H[((int64_t)i * m + i) * T + t];
    const double den = d.x * d.x + d.y * d.y;
    y[(int64_t)i * T + t] = make_double2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den);
  }
}

}  // namespace

extern "C" {

int mg_hess_solve(const void* H, const void* g, int m, int k, int64_t T, void* y, void* stream) {
  if (!H || !g || !y || m < 1 || k < 0 || k > m || T < 0) return MG_ERR_ARG;
  if (k == 0 || T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_hess_solve, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(H), reinterpret_cast<const double2*>(g), m, k, T,
                     reinterpret_cast<double2*>(y));
  MG_HIP(hipGetLastError());
  return MG_OK;
}


int mg_givens(void* H, void* H0, void* cs, void* sn, void* g, const double* hn, const double* bnorm, int m, int j,
              int64_t T, double* est, void* stream) {
  if (!H || !H0 || !cs || !sn || !g || !hn || !bnorm || !est || m < 1 || j < 0 || j >= m || T < 0)
    return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, s, reinterpret_cast<double2*>(H), reinterpret_cast<double2*>(H0),
                     reinterpret_cast<double2*>(cs), reinterpret_cast<double2*>(sn), reinterpret_cast<double2*>(g), hn,
                     bnorm, m, j, T, est);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int64_t mg_mdot_work_elems(int64_t k, int64_t T) { return (k < 0 || T < 0) ? 0 : (int64_t)kChunks * k * T; }

int mg_mdot(const void* V, int64_t sV, int64_t ldv, int k, const void* w, int64_t ldw, int64_t n, int64_t T,
            void* work, void* out, int64_t ldo, void* stream) {
  if (!V || !w || !work || !out || k < 0 || n < 0 || T < 0 || ldv < T || ldw < T || ldo < T) return MG_ERR_ARG;
  if (k == 0 || T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t rows = (n + kChunks - 1) / kChunks;
  hipLaunchKernelGGL(k_mdot_part, dim3(kChunks, (unsigned)((T + 63) / 64)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), sV, ldv, k, reinterpret_cast<const double2*>(w), ldw, n, T,
                     rows, reinterpret_cast<double2*>(work));
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mdot_reduce, dim3((unsigned)((T + 63) / 64), (unsigned)k), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(work), k, T, kChunks, reinterpret_cast<double2*>(out), ldo);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_maxpy(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, double alpha, void* w,
             int64_t ldw, int64_t n, int64_t T, void* stream) {
  if (!V || !h || !w || k < 0 || n < 0 || T < 0 || ldv < T || ldw < T || ldh < T) return MG_ERR_ARG;
  if (k == 0 || n == 0 || T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_maxpy, dim3((unsigned)((n * T + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), sV, ldv, k, reinterpret_cast<const double2*>(h), ldh, alpha,
                     reinterpret_cast<double2*>(w), ldw, n, T);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_maxpy_norm(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, double alpha, void* w,
                  int64_t ldw, int64_t n, int64_t T, void* work, double* norms, void* stream) {
  if (!w || !work || !norms || (k > 0 && (!V || !h)) || k < 0 || n < 0 || T < 0 || ldv < T || ldw < T || ldh < T)
    return MG_ERR_ARG;
  if (T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t rows = (n + kChunks - 1) / kChunks;
  double* part = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_maxpy_norm, dim3(kChunks, (unsigned)((T + 63) / 64)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), sV, ldv, k, reinterpret_cast<const double2*>(h), ldh, alpha,
                     reinterpret_cast<double2*>(w), ldw, n, T, rows, part);
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)((T + 63) / 64)), dim3(256), 0, s, part, T, kChunks, norms);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_maxpy_mdot(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, void* w, int64_t ldw,
                  int64_t n, int64_t T, void* work, void* out, int64_t ldo, void* stream) {
  if (!V || !h || !w || !work || !out || k < 1 || k > kFuseMax || n < 0 || T < 0 || ldv < T || ldw < T ||
      ldh < T || ldo < T)
    return MG_ERR_ARG;
  if (T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t rows = (n + kChunks - 1) / kChunks;
  hipLaunchKernelGGL(k_maxpy_mdot, dim3(kChunks, (unsigned)((T + 63) / 64)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), sV, ldv, k, reinterpret_cast<const double2*>(h), ldh,
                     reinterpret_cast<double2*>(w), ldw, n, T, rows, reinterpret_cast<double2*>(work));
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mdot_reduce, dim3((unsigned)((T + 63) / 64), (unsigned)k), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(work), k, T, kChunks, reinterpret_cast<double2*>(out), ldo);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_colnorm(const void* X, int64_t ldx, int64_t n, int64_t T, void* work, double* norms, void* stream) {
  if (!X || !work || !norms || n < 0 || T < 0 || ldx < T) return MG_ERR_ARG;
  if (T == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t rows = (n + kChunks - 1) / kChunks;
  double* part = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_colnorm_part, dim3(kChunks, (unsigned)((T + 63) / 64)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(X), ldx, n, T, rows, part);
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)((T + 63) / 64)), dim3(256), 0, s, part, T, kChunks, norms);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // extern "C"

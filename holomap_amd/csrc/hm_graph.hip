// Consumers of the graph's CSR on the device: connected components
// (src/components.rs:13-39) and the symmetric-normalised Laplacian
// (src/sparse.rs:55-93), for graphs the fuzzy stage produced (symmetric,
// rows sorted by column).
//
// Components. The reference labels components in first-seen order by a
// search from each unlabelled vertex in ascending order, so a component's
// label is the rank of its smallest vertex among all components' smallest
// vertices. Here every vertex converges to its component's smallest vertex
// by repeated min-propagation over its neighbours with pointer jumping (each
// thread writes only its own label, so a step is race-free in effect: every
// value read is a label of the same component and labels only decrease);
// the host repeats steps until one changes nothing (a fixed point, at most
// diameter + 1 steps), then roots are ranked by an exclusive scan.
//
// Laplacian. degrees d_i = sum of the row's f64-widened weights, folded left
// to right from -0.0 (Rust's f64 Sum); inv_sqrt_i = 1 / sqrt(d_i); the
// normalised adjacency S_ij = w_ij * (inv_sqrt_i * inv_sqrt_j); L x = x - S x
// with each row's S x folded left to right from -0.0. A thread per row keeps
// every fold in the reference's order, so all values are bit-identical.
#include "hm_internal.hpp"

namespace hm {

__global__ void cc_init_kernel(uint32_t* __restrict__ lab, int64_t n) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) lab[v] = (uint32_t)v;
}

// One propagation step: lab[v] = min(lab[v], lab[u] for neighbours u), then
// one jump through the label; *changed = 1 (a plain store: every writer
// stores the same value) if the vertex's label went down.
__global__ void cc_step_kernel(const uint64_t* __restrict__ indptr,
                               const uint32_t* __restrict__ indices, int64_t n,
                               uint32_t* __restrict__ lab, uint32_t* __restrict__ changed) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t cur = lab[v];
  uint32_t m = cur;
  for (uint64_t p = indptr[v]; p < indptr[v + 1]; ++p) {
    const uint32_t l = lab[indices[p]];
    m = l < m ? l : m;
  }
  const uint32_t j = lab[m];
  m = j < m ? j : m;
  if (m < cur) {
    lab[v] = m;
    *changed = 1u;
  }
}

// Pointer jumping: lab[v] = lab[lab[v]].
__global__ void cc_jump_kernel(uint32_t* __restrict__ lab, int64_t n) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) lab[v] = lab[lab[v]];
}

__global__ void cc_root_kernel(const uint32_t* __restrict__ lab, int64_t n,
                               uint32_t* __restrict__ root) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) root[v] = lab[v] == (uint32_t)v ? 1u : 0u;
}

// label[v] = rank of v's root among the roots (rank: exclusive scan of root).
__global__ void cc_label_kernel(const uint32_t* __restrict__ lab,
                                const uint32_t* __restrict__ rank, int64_t n,
                                uint32_t* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[v] = rank[lab[v]];
}

// Degrees and 1 / sqrt(degree), a thread per row (src/sparse.rs:60-67).
__global__ void lap_degree_kernel(const uint64_t* __restrict__ indptr,
                                  const float* __restrict__ vals, int64_t n,
                                  double* __restrict__ deg, double* __restrict__ inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = -0.0;
  for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) d = d + (double)vals[p];
  if (deg) deg[i] = d;
  inv[i] = 1.0 / sqrt(d);
}

// S values, a thread per row (src/sparse.rs:68-73).
__global__ void lap_scale_kernel(const uint64_t* __restrict__ indptr,
                                 const uint32_t* __restrict__ indices,
                                 const float* __restrict__ vals, const double* __restrict__ inv,
                                 int64_t n, double* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = inv[i];
  for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) s[p] = (double)vals[p] * (a * inv[indices[p]]);
}

// y = x - S x (src/sparse.rs:87-93). With svals == nullptr the S values are
// formed on the fly from the f32 weights and inv (the same two products, so
// the same bits).
__global__ void lap_matvec_kernel(const uint64_t* __restrict__ indptr,
                                  const uint32_t* __restrict__ indices,
                                  const double* __restrict__ svals,
                                  const float* __restrict__ vals, const double* __restrict__ inv,
                                  int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = -0.0;
  if (svals) {
    for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) sx = sx + svals[p] * x[indices[p]];
  } else {
    const double a = inv[i];
    for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
      const uint32_t j = indices[p];
      const double sv = (double)vals[p] * (a * inv[j]);
      sx = sx + sv * x[j];
    }
  }
  y[i] = x[i] - sx;
}

// ---- edge schedule of the SGD layout (src/sgd.rs:36-63) ----
// max_w = fold(0.0, f32::max) over the weights (a max: order-free, two
// launches); cutoff = max_w / prune_epochs in f32; an edge is kept unless
// w < cutoff, in COO order (flags, exclusive scan, scatter), with
// epochs_per_sample = f64(max_w) / f64(w).
__device__ __forceinline__ float block_max_f32(float v) {
  __shared__ float red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ __launch_bounds__(256) void sched_max_part_kernel(const float* __restrict__ w,
                                                             int64_t m, float* __restrict__ part) {
  float v = 0.0f;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < m; e += (int64_t)gridDim.x * 256)
    v = fmaxf(v, w[e]);
  const float r = block_max_f32(v);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void sched_max_kernel(const float* __restrict__ part, int np,
                                                        float* __restrict__ out) {
  float v = 0.0f;
  for (int i = threadIdx.x; i < np; i += 256) v = fmaxf(v, part[i]);
  const float r = block_max_f32(v);
  if (threadIdx.x == 0) *out = r;
}

__global__ void sched_flag_kernel(const float* __restrict__ w, int64_t m,
                                  const float* __restrict__ maxw, uint32_t prune_epochs,
                                  uint32_t* __restrict__ keep) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const float cutoff = *maxw / (float)prune_epochs;
  keep[e] = w[e] < cutoff ? 0u : 1u;
}

__global__ void sched_scatter_kernel(const uint32_t* __restrict__ rows,
                                     const uint32_t* __restrict__ cols,
                                     const float* __restrict__ w, int64_t m,
                                     const uint32_t* __restrict__ keep,
                                     const uint32_t* __restrict__ pos,
                                     const float* __restrict__ maxw, uint32_t* __restrict__ head,
                                     uint32_t* __restrict__ tail, double* __restrict__ eps) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m || !keep[e]) return;
  const uint32_t p = pos[e];
  head[p] = rows[e];
  tail[p] = cols[e];
  eps[p] = (double)*maxw / (double)w[e];
}

}  // namespace hm

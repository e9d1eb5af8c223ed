// Evaluation-key generation on the GPU: the same words as the host's generate_bsk /
// generate_ksk (hb_keys.cpp), from the counter-based samplers of hb_sample.hpp.
//
// One thread per coefficient of every RLWE sample draws its ChaCha20 block: the uniform
// a-coefficient goes to the key layout, the Gaussian error (plus the gadget value at X^0)
// to a coefficient-domain buffer. The buffer is transformed by the NTT microbench kernel
// (canonical in and out, the key layout's slot order), and a second pass forms
// b = NTT(e) + a ⊙ s (+ B_ks^i NTT(s_ring) for the key-switching key) and the c_neg_s_m
// rows' a += B^j. A last kernel turns canonical key words into the device form the
// bootstrap kernels read (× N^-1, Montgomery).
#include <hip/hip_runtime.h>

#include "hb_kernels.hpp"
#include "hb_sample.hpp"

namespace hb {

namespace {

// Canonical modular arithmetic (the host keygen's HostNtt add / mul).
__device__ __forceinline__ u64 addq(u64 a, u64 b, u64 q) {
  const u64 s = a + b;
  return (s < a || s >= q) ? s - q : s;
}
__device__ __forceinline__ u64 mulq(u64 a, u64 b, u64 q, int is_goldilocks) {
  if (is_goldilocks) return Gl::canon(Gl::mul(a, b));
  return (a * b) % q;
}

// Row r of the sample grid: BSK row r = (i * 2 + c) * L + j (stream (BSK, i), blocks from
// (c L + j) N); KSK row r (stream (KSK, r), blocks from 0). One thread per coefficient.
__global__ __launch_bounds__(256) void keygen_sample_kernel(KeygenArgs A, u64* __restrict__ keys,
                                                            u64* __restrict__ err) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)A.rows * A.N) return;
  const u32 r = (u32)(t / A.N), k = (u32)(t % A.N);
  u32 stream_row, first;
  if (A.is_bsk) {
    stream_row = r / (2 * A.L);
    first = (r % (2 * A.L)) * A.N;
  } else {
    stream_row = r;
    first = 0;
  }
  u64 a, e;
  key_coefficient(A.key, A.kind, stream_row, first + k, A.q, A.sigma, &a, &e);
  if (A.is_bsk && k == 0) {
    // c = 1 rows of a key bit 1 carry the gadget value B^j at X^0
    const u32 cj = r % (2 * A.L), c = cj / A.L, j = cj % A.L;
    if (c == 1 && A.lwe_sk[stream_row]) e = addq(e, 1ull << (A.basis_bits * j), A.q);
  }
  keys[(size_t)r * 2 * A.N + k] = a;
  err[(size_t)r * A.N + k] = e;
}

__global__ __launch_bounds__(256) void keygen_combine_kernel(KeygenArgs A, u64* __restrict__ keys,
                                                             const u64* __restrict__ err,
                                                             const u64* __restrict__ s_ntt,
                                                             const u64* __restrict__ ring_ntt) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)A.rows * A.N) return;
  const u32 r = (u32)(t / A.N), k = (u32)(t % A.N);
  u64* a = keys + (size_t)r * 2 * A.N;
  const u64 av = a[k];
  u64 b = addq(err[(size_t)r * A.N + k], mulq(av, s_ntt[k], A.q, A.is_goldilocks), A.q);
  if (A.is_bsk) {
    const u32 i = r / (2 * A.L), cj = r % (2 * A.L), c = cj / A.L, j = cj % A.L;
    if (c == 0 && A.lwe_sk[i]) a[k] = addq(av, 1ull << (A.basis_bits * j), A.q);
  } else {
    // b += B_ks^r · NTT(s_ring): B_ks^r mod q by repeated products, as the host's running
    // product s[k] *= B_ks
    u64 sr = ring_ntt[k];
    for (u32 m = 0; m < r; m++) sr = mulq(sr, 1ull << A.basis_bits, A.q, A.is_goldilocks);
    b = addq(b, sr, A.q);
  }
  a[A.N + k] = b;
}

// Canonical key words -> device form (value × scale, Montgomery for the kernels).
__global__ __launch_bounds__(256) void keys_to_device_kernel(const u64* __restrict__ src,
                                                             void* __restrict__ dst, size_t words,
                                                             u64 scale, int is_goldilocks, u64 q) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= words) return;
  if (is_goldilocks) {
    static_cast<u64*>(dst)[t] = Gl::to_mont(Gl::canon(Gl::mul(src[t], scale)));
  } else {
    const u32 x = (u32)((src[t] * scale) % q);
    static_cast<u32*>(dst)[t] = Bb::to_mont(x);
  }
}

dim3 grid_for(size_t threads) { return dim3((unsigned)((threads + 255) / 256)); }

}  // namespace

hipError_t launch_keygen(const KeygenArgs& a, u64* keys, u64* err, const u64* s_ntt,
                         const u64* ring_ntt, const void* tw_fwd, u32 cus, hipStream_t s) {
  const size_t threads = (size_t)a.rows * a.N;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(keygen_sample_kernel, grid_for(threads), dim3(256), 0, s, a, keys, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_ntt(a.is_goldilocks, 0, 10, err, a.rows, tw_fwd, 0, cus, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(keygen_combine_kernel, grid_for(threads), dim3(256), 0, s, a, keys, err,
                     s_ntt, ring_ntt);
  return hipGetLastError();
}

hipError_t launch_keys_to_device(const u64* src, void* dst, size_t words, u64 scale,
                                 int is_goldilocks, u64 q, hipStream_t s) {
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(keys_to_device_kernel, grid_for(words), dim3(256), 0, s, src, dst, words,
                     scale, is_goldilocks, q);
  return hipGetLastError();
}

}  // namespace hb

// Description length of a labelling without any device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).  Every intermediate (the
// partial histogram planes, the dense block matrix, its CSR/CSC, degrees, logs and tile
// sums) lives in one fixed file-scope device array, `scratch`, of NA_SCRATCH_BYTES per
// device; inputs and the result are the caller's device buffers.  The computation is the
// small rebuild path's (B <= 128) followed by the canonical data-term sum and the model
// term, so the result equals gsap_mdl_i32's bit for bit.
//
// The array's host-side name is not a device address: each device's address is resolved
// once with hipGetSymbolAddress (noalloc_init_device, called by gsap_create with the device
// current) and only that address is handed to kernels.  Calls on one device share the
// array, so calls that overlap on one device race; a call leaves the array holding its own
// intermediates.
#include <mutex>
#include <string>

#include "../../include/gsap/gsap.h"
#include "engine.h"

namespace gsap {

constexpr int32_t NA_MAX_B = 128;                     // blocks
constexpr int64_t NA_BB = (int64_t)NA_MAX_B * NA_MAX_B;
constexpr int32_t NA_PLANES = 64;                     // partial histogram planes
constexpr int32_t NA_MAX_DEVICES = 64;

// byte offsets of the scratch regions (each 256-byte aligned)
constexpr int64_t na_align(int64_t x) { return (x + 255) & ~(int64_t)255; }
constexpr int64_t NA_OFF_PLANES = 0;
constexpr int64_t NA_OFF_D = na_align(NA_OFF_PLANES + 4 * NA_PLANES * NA_BB);
constexpr int64_t NA_OFF_ROWPTR = na_align(NA_OFF_D + 4 * NA_BB);
constexpr int64_t NA_OFF_COLPTR = na_align(NA_OFF_ROWPTR + 4 * (NA_MAX_B + 1));
constexpr int64_t NA_OFF_COL = na_align(NA_OFF_COLPTR + 4 * (NA_MAX_B + 1));
constexpr int64_t NA_OFF_VAL = na_align(NA_OFF_COL + 4 * NA_BB);
constexpr int64_t NA_OFF_EROW = na_align(NA_OFF_VAL + 4 * NA_BB);
constexpr int64_t NA_OFF_ROW = na_align(NA_OFF_EROW + 4 * NA_BB);
constexpr int64_t NA_OFF_CVAL = na_align(NA_OFF_ROW + 4 * NA_BB);
constexpr int64_t NA_OFF_DOUT = na_align(NA_OFF_CVAL + 4 * NA_BB);
constexpr int64_t NA_OFF_DIN = na_align(NA_OFF_DOUT + 4 * NA_MAX_B);
constexpr int64_t NA_OFF_LDOUT = na_align(NA_OFF_DIN + 4 * NA_MAX_B);
constexpr int64_t NA_OFF_LDIN = na_align(NA_OFF_LDOUT + 8 * NA_MAX_B);
constexpr int64_t NA_OFF_TILES = na_align(NA_OFF_LDIN + 8 * NA_MAX_B);
constexpr int64_t NA_OFF_DATA = na_align(NA_OFF_TILES + 8 * (NA_BB / 1024 + 1));
constexpr int64_t NA_OFF_ERR = na_align(NA_OFF_DATA + 8);
constexpr int64_t NA_SCRATCH_BYTES = na_align(NA_OFF_ERR + 4);  // about 4.5 MiB

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

static std::mutex na_mu;
static unsigned char *na_ptr[NA_MAX_DEVICES] = {};

bool noalloc_init_device(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return false;
  std::lock_guard<std::mutex> lk(na_mu);
  if (na_ptr[device]) return true;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
  na_ptr[device] = static_cast<unsigned char *>(p);
  return true;
}

static unsigned char *noalloc_ptr(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(na_mu);
  return na_ptr[device];
}

__global__ void k_na_clear(uint32_t *err) { *err = 0; }

// Histogram of the labelling's block matrix, B <= 128, whole B x B counter set in LDS per
// workgroup: workgroup c takes vertices [N c / C, N (c + 1) / C), a wave per vertex, and
// writes plane c.  Entries whose ids, weight or row pointers are out of range are skipped
// and counted in *err.
__global__ __launch_bounds__(512) void k_na_hist(const int32_t *optr, const int32_t *oadj,
                                                 const int32_t *ow, const int32_t *lab,
                                                 int64_t N, int64_t E, int32_t B, int32_t C,
                                                 int32_t *planes, uint32_t *err) {
  extern __shared__ int32_t T[];
  const int32_t nt = B * B;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) T[i] = 0;
  __syncthreads();
  const int64_t c = blockIdx.x;
  const int64_t v0 = N * c / C, v1 = N * (c + 1) / C;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t bad = 0;
  for (int64_t v = v0 + w; v < v1; v += nw) {
    const int32_t r = lab[v];
    const int64_t o0 = optr[v], o1 = optr[v + 1];
    if (r < 0 || r >= B || o0 < 0 || o1 < o0 || o1 > E) {
      bad += lane == 0;
      continue;
    }
    for (int64_t j = o0 + lane; j < o1; j += 64) {
      const int32_t u = oadj[j], x = ow[j];
      const int32_t t = (u >= 0 && u < N) ? lab[u] : -1;
      if (t < 0 || t >= B || x <= 0) {
        ++bad;
        continue;
      }
      atomicAdd(&T[r * B + t], x);
    }
  }
  if (bad) atomicAdd(err, bad);
  __syncthreads();
  int32_t *out = planes + c * (int64_t)nt;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) out[i] = T[i];
}

// S = model term - data term (the library's model_S on the device, same operations and
// canonical log2, so the same bits); NaN when the histogram skipped an entry
__global__ void k_na_model(const double *data, const uint32_t *err, int64_t N, int64_t E,
                           int32_t B, int32_t model_kind, double *S) {
  if (threadIdx.x != 0) return;
  const double LN2 = 0.6931471805599453094;
  const double b = (double)B, dE = (double)E;
  const double x = (b * b) / dE;
  double model;
  if (model_kind == GSAP_MODEL_CORRECTED)
    model = dE * ((1.0 + x) * (clog2(1.0 + x) * LN2) - x * (clog2(x) * LN2)) +
            (double)N * (clog2(b) * LN2);
  else
    model = dE * (1.0 + x) * (clog2(1.0 + x) * LN2) - x * (clog2(x) * LN2) +
            (double)N * (clog2(b) * LN2);
  *S = *err ? __builtin_nan("") : model - *data;
}

int noalloc_mdl(int device, const int32_t *optr, const int32_t *oadj, const int32_t *ow,
                const int32_t *lab, int64_t N, int64_t E, int32_t B, int32_t model_kind,
                double *S, hipStream_t st, std::string &msg) {
  if (!optr || !oadj || !ow || !lab || !S) {
    msg = "gsap_mdl_noalloc_i32: NULL argument";
    return GSAP_ERR_INVALID;
  }
  if (B < 1 || B > NA_MAX_B) {
    msg = "gsap_mdl_noalloc_i32: B = " + std::to_string(B) + " is outside 1.." +
          std::to_string(NA_MAX_B) + " (the fixed scratch array's limit)";
    return GSAP_ERR_INVALID;
  }
  if (N < 1 || N > INT32_MAX || E < 1 || E > INT32_MAX) {
    msg = "gsap_mdl_noalloc_i32: N and E must be in 1..2^31-1";
    return GSAP_ERR_INVALID;
  }
  if (model_kind != GSAP_MODEL_REFERENCE && model_kind != GSAP_MODEL_CORRECTED) {
    msg = "gsap_mdl_noalloc_i32: unknown model";
    return GSAP_ERR_INVALID;
  }
  unsigned char *p = noalloc_ptr(device);
  if (!p) {
    msg = "gsap_mdl_noalloc_i32: scratch array of device " + std::to_string(device) +
          " not resolved";
    return GSAP_ERR_INTERNAL;
  }
  auto at = [&](int64_t off) { return p + off; };
  int32_t *planes = reinterpret_cast<int32_t *>(at(NA_OFF_PLANES));
  int32_t *D = reinterpret_cast<int32_t *>(at(NA_OFF_D));
  int32_t *rowptr = reinterpret_cast<int32_t *>(at(NA_OFF_ROWPTR));
  int32_t *colptr = reinterpret_cast<int32_t *>(at(NA_OFF_COLPTR));
  int32_t *col = reinterpret_cast<int32_t *>(at(NA_OFF_COL));
  int32_t *val = reinterpret_cast<int32_t *>(at(NA_OFF_VAL));
  int32_t *erow = reinterpret_cast<int32_t *>(at(NA_OFF_EROW));
  int32_t *row = reinterpret_cast<int32_t *>(at(NA_OFF_ROW));
  int32_t *cval = reinterpret_cast<int32_t *>(at(NA_OFF_CVAL));
  int32_t *dout = reinterpret_cast<int32_t *>(at(NA_OFF_DOUT));
  int32_t *din = reinterpret_cast<int32_t *>(at(NA_OFF_DIN));
  double *ldout = reinterpret_cast<double *>(at(NA_OFF_LDOUT));
  double *ldin = reinterpret_cast<double *>(at(NA_OFF_LDIN));
  double *tiles = reinterpret_cast<double *>(at(NA_OFF_TILES));
  double *data = reinterpret_cast<double *>(at(NA_OFF_DATA));
  uint32_t *err = reinterpret_cast<uint32_t *>(at(NA_OFF_ERR));
  // about 64k edges per workgroup, at most NA_PLANES planes
  int32_t C = (int32_t)std::min<int64_t>(NA_PLANES, (E + 65535) / 65536);
  if (C > N) C = (int32_t)N;
  if (C < 1) C = 1;
  const int64_t BB = (int64_t)B * B;
  k_na_clear<<<1, 1, 0, st>>>(err);
  k_na_hist<<<C, 512, (size_t)BB * 4, st>>>(optr, oadj, ow, lab, N, E, B, C, planes, err);
  launch_small_from_planes(planes, C, B, D, rowptr, col, val, erow, colptr, row, cval, dout, din,
                           ldout, ldin, st);
  BMView m;
  m.B = B;
  m.rowptr = rowptr; m.col = col; m.val = val;
  m.colptr = colptr; m.row = row; m.cval = cval;
  m.dout = dout; m.din = din; m.ldout = ldout; m.ldin = ldin;
  m.dense = nullptr; m.gcol = nullptr; m.grow = nullptr;
  launch_mdl_sum(m, erow, BB, tiles, data, nullptr, 0, st);
  k_na_model<<<1, 64, 0, st>>>(data, err, N, E, B, model_kind, S);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    msg = std::string("gsap_mdl_noalloc_i32: ") + hipGetErrorString(e);
    return GSAP_ERR_DEVICE;
  }
  return GSAP_OK;
}

}  // namespace gsap

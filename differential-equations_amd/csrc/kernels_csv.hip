// Solution::to_csv (src/solution.rs:190-230) for every trajectory of an
// ensemble solve, on the GPU.
//
// Work is cut into slots: slot r of trajectory i is its header (r = 0) or
// its point r - 1.  Three passes over a chunk of trajectories:
//   csv_points_kernel  the number of points of each trajectory and, for t_eval
//                      output, where they sit in t_eval;
//   csv_len_kernel     one thread per slot: the byte length of the slot's line
//                      (each number through fmt_prepare: Ryu digits, layout);
//   exclusive scan     (hipCUB) of the lengths: every slot's output position,
//                      trajectory i's document starting at its header's;
//   csv_write_kernel   256 consecutive slots per block: each thread formats
//                      its line into LDS at its position, then the block copies
//                      its contiguous byte range out with aligned 4-byte stores
//                      (byte stores only at the range's ragged ends).
// The numbers are formatted twice (length, then text): a double's digits cost
// less to recompute than to store and reload.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "de_fmt.h"
#include "de_launch.h"

namespace de {
namespace {

constexpr int kCsvBlock = 256;
// Slots per chunk: bounds the temporary arrays (12 bytes per slot).
constexpr int64_t kCsvChunkSlots = int64_t(1) << 26;

DE_DEV int header_len(int dim) {
  int n = 2;  // "t" and "\n"
  for (int c = 0; c < dim; ++c) n += 2 + (c >= 10 ? 2 : 1);
  return n;
}

DE_DEV int header_emit(int dim, char* dst) {
  int n = 0;
  dst[n++] = 't';
  for (int c = 0; c < dim; ++c) {
    dst[n++] = ',';
    dst[n++] = 'y';
    if (c >= 10) dst[n++] = (char)('0' + c / 10);
    dst[n++] = (char)('0' + c % 10);
  }
  dst[n++] = '\n';
  return n;
}

// Point p of trajectory i: its time and state.
DE_DEV void csv_point(const CsvArgs& a, int64_t i, int32_t p, int32_t head, int32_t run, double* t,
                      double* y) {
  const int64_t n = a.n;
  const int dim = a.dim;
  if (a.source == kCsvFinal) {
    *t = a.t_final[i];
    for (int c = 0; c < dim; ++c) y[c] = a.y_final[c * n + i];
  } else if (a.source == kCsvEvents) {
    *t = a.event_t[(int64_t)p * n + i];
    for (int c = 0; c < dim; ++c) y[c] = a.event_y[((int64_t)p * dim + c) * n + i];
  } else {
    const int64_t k = p < head ? 0 : (int64_t)run + p - head;
    *t = a.t_eval[k];
    for (int c = 0; c < dim; ++c) y[c] = a.dense[(k * dim + c) * n + i];
  }
}

// The points of each trajectory of the chunk.  t_eval output: TEvalSolout's
// first call (t_prev = t_curr = t0) produces index 0 if it equals t0 and
// passes over every other point not beyond t0; every later point is produced
// in order (src/solout/t_eval.rs:100-135).  So the points are index 0 when
// `head`, then the run from index `run`.
__global__ void csv_points_kernel(CsvArgs a, int64_t i0, int64_t cnt, int32_t* count,
                                  int32_t* head, int32_t* run) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t i = i0 + j;
  int32_t c = 0, h = 0, r = 0;
  if (a.source == kCsvFinal) {
    c = 1;
  } else if (a.source == kCsvEvents) {
    c = a.n_events ? a.n_events[i] : 0;
  } else {
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    const double tf = a.tf ? a.tf[i] : a.tfs;
    const bool fwd = tf > t0;
    h = a.n_eval > 0 && a.t_eval[0] == t0;
    int32_t idx = 0;
    while (idx < a.n_eval) {
      const double te = a.t_eval[idx];
      if ((fwd && te > t0) || (!fwd && te < t0)) break;
      ++idx;
    }
    r = idx;
    c = a.n_dense ? a.n_dense[i] : 0;
  }
  // Only the stored points: max_events of them, and the t_eval rows that exist.
  const int32_t cap = a.source == kCsvTeval ? h + (a.n_eval - r) : a.max_rows;
  count[j] = c < 0 ? 0 : (c > cap ? cap : c);
  head[j] = h;
  run[j] = r;
}

// Byte length of slot s of the chunk (0 past a trajectory's points);
// len[nslots] = 0 so that the scan's last entry is the chunk's total.
__global__ void csv_len_kernel(CsvArgs a, int64_t i0, uint32_t nslots, uint32_t per,
                               const int32_t* count, const int32_t* head, const int32_t* run,
                               int64_t* len) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslots) return;
  if (s == nslots) {
    len[s] = 0;
    return;
  }
  const uint32_t j = s / per, r = s - j * per;
  const int32_t cnt = count[j];
  int32_t bytes = 0;
  if (cnt > 0) {
    if (r == 0) {
      bytes = header_len(a.dim);
    } else if ((int32_t)r - 1 < cnt) {
      double t, y[kCsvMaxDim];
      csv_point(a, i0 + j, (int32_t)r - 1, head[j], run[j], &t, y);
      bytes = fmt_prepare(t).len + a.dim + 1;
      for (int c = 0; c < a.dim; ++c) bytes += fmt_prepare(y[c]).len;
    }
  }
  len[s] = bytes;
}

// offsets[i0 + j] = base + pos[j * per]; the chunk's total is added to base.
__global__ void csv_offsets_kernel(int64_t cnt, uint32_t per, uint32_t nslots, const int64_t* pos,
                                   int64_t* base, int64_t* offsets, int64_t i0, int last) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = *base;
  if (j < cnt) offsets[i0 + j] = b + pos[j * per];
  if (j == cnt && last) offsets[i0 + cnt] = b + pos[nslots];
}

__global__ void csv_base_kernel(uint32_t nslots, const int64_t* pos, int64_t* base) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *base += pos[nslots];
}

__global__ void csv_zero_kernel(int64_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

__global__ __launch_bounds__(kCsvBlock) void csv_write_kernel(
    CsvArgs a, int64_t i0, uint32_t nslots, uint32_t per, const int32_t* count,
    const int32_t* head, const int32_t* run, const int64_t* pos, const int64_t* base, char* out,
    int64_t cap) {
  extern __shared__ char lds[];
  const uint32_t s0 = blockIdx.x * kCsvBlock;
  const uint32_t s = s0 + threadIdx.x;
  const uint32_t s1 = s0 + kCsvBlock < nslots ? s0 + kCsvBlock : nslots;
  const int64_t b = *base;
  // A block whose bytes would pass the caller's buffer writes nothing (the
  // host then reports the size mismatch).
  if (b + pos[s1] > cap) return;
  // The block's bytes [g_lo, g_hi) of out, staged in LDS from the 4-byte
  // aligned address at or below g_lo.
  const uintptr_t g_lo = (uintptr_t)(out + b + pos[s0]);
  const uintptr_t g_hi = (uintptr_t)(out + b + pos[s1]);
  const uintptr_t g_al = g_lo & ~(uintptr_t)3;
  if (s < nslots) {
    const int64_t p0 = pos[s], bytes = pos[s + 1] - p0;
    if (bytes > 0) {
      char* dst = lds + ((uintptr_t)(out + b + p0) - g_al);
      const uint32_t j = s / per, r = s - j * per;
      if (r == 0) {
        header_emit(a.dim, dst);
      } else {
        double t, y[kCsvMaxDim];
        csv_point(a, i0 + j, (int32_t)r - 1, head[j], run[j], &t, y);
        FmtF64 f = fmt_prepare(t);
        fmt_emit(f, dst);
        int k = f.len;
        for (int c = 0; c < a.dim; ++c) {
          dst[k++] = ',';
          f = fmt_prepare(y[c]);
          fmt_emit(f, dst + k);
          k += f.len;
        }
        dst[k] = '\n';
      }
    }
  }
  __syncthreads();
  const uintptr_t w_lo = (g_lo + 3) & ~(uintptr_t)3;  // first whole word
  const uintptr_t w_hi = g_hi & ~(uintptr_t)3;        // end of the whole words
  if (w_lo >= w_hi) {
    for (uintptr_t g = g_lo + threadIdx.x; g < g_hi; g += kCsvBlock)
      *(char*)g = lds[g - g_al];
    return;
  }
  if (threadIdx.x < w_lo - g_lo) *(char*)(g_lo + threadIdx.x) = lds[g_lo + threadIdx.x - g_al];
  if (threadIdx.x < g_hi - w_hi) *(char*)(w_hi + threadIdx.x) = lds[w_hi + threadIdx.x - g_al];
  for (uintptr_t g = w_lo + 4 * (uintptr_t)threadIdx.x; g < w_hi; g += 4 * kCsvBlock)
    *(uint32_t*)g = *(const uint32_t*)(lds + (g - g_al));
}

}  // namespace

// Sizes (offsets) and, when out is not null, the documents of every
// trajectory into out[0 .. cap); device pointers.  Returns 0, or -1 on a HIP
// error.
int launch_csv(const CsvArgs& a, int64_t* offsets, char* out, int64_t cap, hipStream_t s) {
  if (a.n == 0) {
    hipLaunchKernelGGL(csv_zero_kernel, dim3(1), dim3(1), 0, s, offsets);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  const int64_t per = (int64_t)a.max_rows + 1;
  int64_t chunk = kCsvChunkSlots / per;
  if (chunk < 1) chunk = 1;
  if (chunk > a.n) chunk = a.n;
  const int64_t max_slots = chunk * per;
  int32_t* cnt_head_run = nullptr;
  int64_t* len = nullptr;
  int64_t* pos = nullptr;
  int64_t* base = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int64_t*)nullptr,
                                                  (int64_t*)nullptr, (int)(max_slots + 1), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&cnt_head_run, sizeof(int32_t) * 3 * chunk, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&len, sizeof(int64_t) * (max_slots + 1), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&pos, sizeof(int64_t) * (max_slots + 1), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&base, sizeof(int64_t), s);
  if (e == hipSuccess) e = hipMallocAsync(&scan_tmp, scan_bytes ? scan_bytes : 1, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(csv_zero_kernel, dim3(1), dim3(1), 0, s, base);
    e = hipGetLastError();
  }
  const size_t lds = (size_t)kCsvBlock * csv_max_slot_bytes(a.dim) + 8;
  for (int64_t i0 = 0; e == hipSuccess && i0 < a.n; i0 += chunk) {
    const int64_t cnt = a.n - i0 < chunk ? a.n - i0 : chunk;
    const uint32_t nslots = (uint32_t)(cnt * per);
    int32_t* count = cnt_head_run;
    int32_t* head = cnt_head_run + chunk;
    int32_t* run = cnt_head_run + 2 * chunk;
    hipLaunchKernelGGL(csv_points_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a,
                       i0, cnt, count, head, run);
    hipLaunchKernelGGL(csv_len_kernel, dim3((nslots + 1 + 255) / 256), dim3(256), 0, s, a, i0,
                       nslots, (uint32_t)per, (const int32_t*)count, (const int32_t*)head,
                       (const int32_t*)run, len);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, len, pos, (int)(nslots + 1), s);
    if (e != hipSuccess) break;
    const int last = i0 + cnt == a.n;
    hipLaunchKernelGGL(csv_offsets_kernel, dim3((unsigned)((cnt + 1 + 255) / 256)), dim3(256), 0,
                       s, cnt, (uint32_t)per, nslots, (const int64_t*)pos, base, offsets, i0, last);
    if (out)
      hipLaunchKernelGGL(csv_write_kernel, dim3((nslots + kCsvBlock - 1) / kCsvBlock),
                         dim3(kCsvBlock), lds, s, a, i0, nslots, (uint32_t)per,
                         (const int32_t*)count, (const int32_t*)head, (const int32_t*)run,
                         (const int64_t*)pos, (const int64_t*)base, out, cap);
    hipLaunchKernelGGL(csv_base_kernel, dim3(1), dim3(1), 0, s, nslots, (const int64_t*)pos, base);
    e = hipGetLastError();
  }
  for (void* p : {(void*)cnt_head_run, (void*)len, (void*)pos, (void*)base, scan_tmp}) {
    const hipError_t f = hipFreeAsync(p, s);
    if (e == hipSuccess) e = f;
  }
  return e == hipSuccess ? 0 : -1;
}

}  // namespace de

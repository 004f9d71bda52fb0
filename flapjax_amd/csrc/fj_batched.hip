// Fused multi-case pair sums for the serving loop: one launch evaluates the
// induced velocity of many independent cases, each case's targets seeing
// only its own filaments.
//
// A task table (device memory, FjTask records) gives per case the offsets of
// its targets / filaments / outputs in packed buffers and its tiling; the
// grid is a fixed number of workgroups that stride over the work items of
// all tasks (work item -> task by a search over the tasks' prefix offsets),
// so one launch shape serves any mix of case sizes and the launch can be
// replayed (e.g. from a captured graph) while the table contents change.
// Partial sums go to a per-task region of a packed partial buffer; the
// reduction pass applies each task's epilogue (store / add / convect).
//
// FP64 tasks whose filaments are whole lattices (the serving tick's) run the
// lattice kernel of fj_lattice.hip instead: a work item is a 64-target tile
// and FJ_BS_WAVES (strip, column block) lattice tasks, one per wave, over
// strip-column records built for all tasks by one launch.
#include "fj_common.hpp"
#include "fj_lattice_dev.hpp"


struct FjTask {
  const double* tgt;       // this task's targets (n_tgt, 3)
  double* out;             // this task's outputs (n_tgt, 3)
  int tgt_off, n_tgt;      // first row in the tasks' concatenated target list
  int seg_off, n_seg;      // filament rows in the packed table
  int tiles, chunks, chunk_len;
  int work_off;            // first work item of this task
  int part_off;            // first double of this task's partials
  int ld;                  // padded target count (tiles * 64)
  int mode;                // FJ_OUT_*
  double dt, time;
  FjFlow flow;
  // lattice form: lattices (table row, rings m x n, column blocks per strip,
  // first lattice task, first strip-column record), column-block width,
  // lattice tasks, records
  int n_lat;
  int lat_off[FJ_TASK_MAX_LAT], lat_m[FJ_TASK_MAX_LAT], lat_n[FJ_TASK_MAX_LAT];
  int lat_cb[FJ_TASK_MAX_LAT], lat_task[FJ_TASK_MAX_LAT], lat_rec[FJ_TASK_MAX_LAT];
  int cb_width, n_ltasks, rec_off, n_recs;
};

__device__ __forceinline__ int fj_find_task(const FjTask* __restrict__ tasks, int n_tasks, int w) {
  // last task whose work_off <= w (tasks sorted by work_off)
  int lo = 0, hi = n_tasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].work_off <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// MIXED: the mixed form of the single-case kernels (FP32 far field, FP64
// near field, FP32 partial sums folded into FP64 every FJ_MIXED_FLUSH
// filaments of a wave)
template <int KIND, bool MIXED>
__global__ __launch_bounds__(FJ_BS_THREADS) void k_batched_partial(
    const FjTask* __restrict__ tasks, const int* __restrict__ counts, const double* __restrict__ seg,
    double* __restrict__ part, FjKernelParams kp) {
  __shared__ double red[FJ_BS_WAVES - 1][3][64];
  const int n_tasks = counts[0];
  const int total = counts[1];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float eps2f = (float)kp.eps2;
  const float rcut2f = (float)(kp.rcut * kp.rcut);
  for (int w = blockIdx.x; w < total; w += gridDim.x) {
    const int ti = __builtin_amdgcn_readfirstlane(fj_find_task(tasks, n_tasks, w));
    const FjTask& T = tasks[ti];
    const int local = w - T.work_off;
    const int tile = local % T.tiles;
    const int chunk = local / T.tiles;
    int t = tile * 64 + lane;
    const int tc = (t < T.n_tgt) ? t : (T.n_tgt - 1);
    const double* xp = T.tgt + 3 * (size_t)tc;
    const double x = xp[0], y = xp[1], z = xp[2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    const int s_begin = chunk * T.chunk_len;
    const int s_end = min(T.n_seg, s_begin + T.chunk_len);
    const double* sb = seg + (size_t)T.seg_off * FJ_SEG_STRIDE;
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    int since_flush = 0;
    for (int s = s_begin + wave; s < s_end; s += FJ_BS_WAVES) {
      const double* sp = sb + (size_t)s * FJ_SEG_STRIDE;
      double rec[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) rec[k] = sp[k];
      if (MIXED) {
        const float* spf = (const float*)(sp + 12);
        float sf[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) sf[k] = spf[k];
        const float r0sq_f = (float)rec[11];
        const float r1x = (float)(x - rec[0]), r1y = (float)(y - rec[1]), r1z = (float)(z - rec[2]);
        const float s1 = r1x * r1x + r1y * r1y + r1z * r1z;
        if (s1 < FJ_NEAR_FACTOR2 * r0sq_f) {
          double vx, vy, vz;
          fj_bs_eval<KIND>(x, y, z, rec, kp, vx, vy, vz);
          ax += vx; ay += vy; az += vz;
        } else {
          float vx, vy, vz;
          fj_bs_eval_f32<KIND>(r1x, r1y, r1z, s1, sf, r0sq_f, eps2f, rcut2f, vx, vy, vz);
          bx += vx; by += vy; bz += vz;
        }
        if (++since_flush == FJ_MIXED_FLUSH) {
          since_flush = 0;
          ax += (double)bx; ay += (double)by; az += (double)bz;
          bx = 0.0f; by = 0.0f; bz = 0.0f;
        }
      } else {
        double vx, vy, vz;
        fj_bs_eval<KIND>(x, y, z, rec, kp, vx, vy, vz);
        ax += vx; ay += vy; az += vz;
      }
    }
    if (MIXED) {
      ax += (double)bx; ay += (double)by; az += (double)bz;
    }
    if (wave > 0) {
      red[wave - 1][0][lane] = ax;
      red[wave - 1][1][lane] = ay;
      red[wave - 1][2][lane] = az;
    }
    __syncthreads();
    if (wave == 0) {
      for (int q = 0; q < FJ_BS_WAVES - 1; ++q) {
        ax += red[q][0][lane]; ay += red[q][1][lane]; az += red[q][2][lane];
      }
      double* o = part + T.part_off + (size_t)chunk * 3 * T.ld;
      o[t] = ax;
      o[T.ld + t] = ay;
      o[2 * T.ld + t] = az;
    }
    __syncthreads();
  }
}

// strip-column records of every lattice task: one thread per record word
__global__ void k_batched_lattice_records(const FjTask* __restrict__ tasks,
                                          const int* __restrict__ counts,
                                          const double* __restrict__ seg,
                                          double* __restrict__ tab, int kernel) {
  const int n_tasks = counts[0];
  const long long total = (long long)counts[3] * FJ_LAT_REC;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < total;
       w += (long long)gridDim.x * blockDim.x) {
    const int q = (int)(w / FJ_LAT_REC), f = (int)(w - (long long)q * FJ_LAT_REC);
    // last task whose rec_off <= q among tasks with records
    int lo = 0, hi = n_tasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].rec_off <= q) lo = mid; else hi = mid - 1;
    }
    const FjTask& T = tasks[lo];
    const int li = (T.n_lat > 1 && q >= T.lat_rec[1]) ? 1 : 0;
    const int local = q - T.lat_rec[li];
    const int n = T.lat_n[li];
    const int s = local / (n + 1), j = local - s * (n + 1);
    tab[w] = fj_lat_record_word(seg, T.lat_off[li], T.lat_m[li], n, s, j, f, kernel);
  }
}

template <int KIND>
__global__ __launch_bounds__(FJ_BS_THREADS, KIND == FJ_KERNEL_CUTOFF ? 4 : 5) void
k_batched_lattice_partial(const FjTask* __restrict__ tasks, const int* __restrict__ counts,
                          const double* __restrict__ tab, double* __restrict__ part,
                          FjKernelParams kp) {
  __shared__ double red[FJ_BS_WAVES - 1][3][64];
  const int n_tasks = counts[0];
  const int total = counts[1];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  const double rcut2 = kp.rcut * kp.rcut;
  for (int w = blockIdx.x; w < total; w += gridDim.x) {
    const int ti = __builtin_amdgcn_readfirstlane(fj_find_task(tasks, n_tasks, w));
    const FjTask& T = tasks[ti];
    const int local = w - T.work_off;
    const int tile = local % T.tiles;
    const int chunk = local / T.tiles;
    const int t = tile * 64 + lane;
    const int tc = (t < T.n_tgt) ? t : (T.n_tgt - 1);
    const double* xp = T.tgt + 3 * (size_t)tc;
    const double x = xp[0], y = xp[1], z = xp[2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    const int lt = chunk * FJ_BS_WAVES + wave;
    if (lt < T.n_ltasks) {
      const int li = (T.n_lat > 1 && lt >= T.lat_task[1]) ? 1 : 0;
      const int m = T.lat_m[li], n = T.lat_n[li];
      const int l2 = lt - T.lat_task[li];
      const int s = l2 / T.lat_cb[li], b = l2 - s * T.lat_cb[li];
      const int i0 = s * FJ_LAT_ROWS;
      const int rows = min(FJ_LAT_ROWS, m - i0);
      const int last = (i0 + rows == m) ? 1 : 0;
      const int jb = b * T.cb_width;
      const int je = min(n + 1, jb + T.cb_width);
      const double* col = tab + (size_t)(T.lat_rec[li] + s * (n + 1)) * FJ_LAT_REC;
      fj_lat_strip<KIND>(col, rows, last, jb, je, x, y, z, e2, rcut2, ax, ay, az);
    }
    if (wave > 0) {
      red[wave - 1][0][lane] = ax;
      red[wave - 1][1][lane] = ay;
      red[wave - 1][2][lane] = az;
    }
    __syncthreads();
    if (wave == 0) {
      for (int q = 0; q < FJ_BS_WAVES - 1; ++q) {
        ax += red[q][0][lane]; ay += red[q][1][lane]; az += red[q][2][lane];
      }
      double* o = part + T.part_off + (size_t)chunk * 3 * T.ld;
      o[t] = ax;
      o[T.ld + t] = ay;
      o[2 * T.ld + t] = az;
    }
    __syncthreads();
  }
}

// one workgroup per 64 packed targets (a block may straddle tasks: each lane
// follows its own task); wave w sums chunks w, w + FJ_BRED_WAVES, ... and
// wave 0 adds the wave sums in order (deterministic for given sizes)
#define FJ_BRED_WAVES 8
__global__ __launch_bounds__(64 * FJ_BRED_WAVES) void k_batched_reduce(
    const FjTask* __restrict__ tasks, const int* __restrict__ counts,
    const double* __restrict__ part) {
  __shared__ double red[FJ_BRED_WAVES - 1][3][64];
  const int n_tasks = counts[0];
  const int n_tgt_total = counts[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.x * 64 + lane;
  // task owning packed target g (tasks sorted by tgt_off)
  int lo = 0;
  if (g < n_tgt_total) {
    int hi = n_tasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].tgt_off <= g) lo = mid; else hi = mid - 1;
    }
  }
  const FjTask& T = tasks[lo];
  const int t = g - T.tgt_off;
  const bool live = g < n_tgt_total && t < T.n_tgt;
  double vx = 0.0, vy = 0.0, vz = 0.0;
  if (live) {
    for (int c = wave; c < T.chunks; c += FJ_BRED_WAVES) {
      const double* o = part + T.part_off + (size_t)c * 3 * T.ld;
      vx += o[t]; vy += o[T.ld + t]; vz += o[2 * T.ld + t];
    }
  }
  if (wave > 0) {
    red[wave - 1][0][lane] = vx;
    red[wave - 1][1][lane] = vy;
    red[wave - 1][2][lane] = vz;
  }
  __syncthreads();
  if (wave != 0 || !live) return;
#pragma unroll
  for (int w = 0; w < FJ_BRED_WAVES - 1; ++w) {
    vx += red[w][0][lane];
    vy += red[w][1][lane];
    vz += red[w][2][lane];
  }
  vx *= FJ_INV_4PI; vy *= FJ_INV_4PI; vz *= FJ_INV_4PI;
  double* op = T.out + 3 * (size_t)t;
  if (T.mode == FJ_OUT_STORE) {
    op[0] = vx; op[1] = vy; op[2] = vz;
  } else if (T.mode == FJ_OUT_ADD) {
    op[0] += vx; op[1] += vy; op[2] += vz;
  } else {
    const double* xp = T.tgt + 3 * (size_t)t;
    double u[3];
    fj_flow_eval(T.flow, xp[0], xp[1], xp[2], T.time, u);
    op[0] = xp[0] + T.dt * (u[0] + vx);
    op[1] = xp[1] + T.dt * (u[1] + vy);
    op[2] = xp[2] + T.dt * (u[2] + vz);
  }
}

// ---------------------------------------------------------------------------
// host side: the caller passes per-case descriptors in HOST memory; the
// library fills the device task table (one async copy) and launches.
size_t fj_batched_desc_bytes(int n_tasks) { return (size_t)n_tasks * sizeof(FjTaskDesc); }

// device task table: the task records and three counters; the host staging
// area is the same size plus room for the task descriptors
extern "C" size_t fj_batched_task_bytes(int n_tasks) {
  return (size_t)n_tasks * sizeof(FjTask) + 64 + fj_batched_desc_bytes(n_tasks);
}

// a task's tiling depends on its own sizes only, so a case's arithmetic (and
// result) is the same whatever else shares the launch: filament chunks of
// 256, but at most about FJ_BS_TARGET_BLOCKS work items per task (large
// tasks take longer chunks, which bounds their partial-sum buffer)
static void fj_task_tiling(int n_tgt, int n_seg, int& tiles, int& chunks, int& chunk_len) {
  tiles = n_tgt > 0 ? fj_cdiv(n_tgt, 64) : 0;  // no work for an empty target set
  chunks = n_seg > 0 ? fj_cdiv(n_seg, 256) : 1;
  const int cap = tiles > 0 ? fj_cdiv(FJ_BS_TARGET_BLOCKS, tiles) : 1;
  if (chunks > cap) chunks = cap;
  chunk_len = n_seg > 0 ? fj_cdiv(n_seg, chunks) : 0;
}

// lattice form of a task: column blocks of 16 chordwise-filament columns
// (about 64 filaments per wave task), widened until the task has at most
// about FJ_BS_TARGET_BLOCKS work items; depends on the task's own sizes only
#ifndef FJ_BATCH_LAT_WIDTH
#define FJ_BATCH_LAT_WIDTH 16
#endif
struct FjLatTiling {
  int tiles, chunks, width, n_ltasks, n_recs;
  int cb[FJ_TASK_MAX_LAT], ltask[FJ_TASK_MAX_LAT], rec[FJ_TASK_MAX_LAT];
};

static FjLatTiling fj_task_lattice_tiling(const FjTaskDesc& d) {
  FjLatTiling L;
  L.tiles = d.n_tgt > 0 ? fj_cdiv(d.n_tgt, 64) : 0;
  const int n_lat = d.n_tgt > 0 ? d.n_lat : 0;  // no targets: no records, no work
  for (int k = 0; k < FJ_TASK_MAX_LAT; ++k) L.cb[k] = L.ltask[k] = L.rec[k] = 0;
  int max_cols = 1;
  for (int k = 0; k < n_lat; ++k) max_cols = d.lat_n[k] + 1 > max_cols ? d.lat_n[k] + 1 : max_cols;
  L.width = max_cols < FJ_BATCH_LAT_WIDTH ? max_cols : FJ_BATCH_LAT_WIDTH;
  for (;;) {
    L.n_ltasks = 0;
    L.n_recs = 0;
    for (int k = 0; k < n_lat; ++k) {
      const bool work = d.lat_m[k] > 0 && d.lat_n[k] > 0;
      L.cb[k] = work ? fj_cdiv(d.lat_n[k] + 1, L.width) : 0;
      L.ltask[k] = L.n_ltasks;
      L.rec[k] = L.n_recs;
      if (work) {
        L.n_ltasks += fj_cdiv(d.lat_m[k], FJ_LAT_ROWS) * L.cb[k];
        L.n_recs += fj_cdiv(d.lat_m[k], FJ_LAT_ROWS) * (d.lat_n[k] + 1);
      }
    }
    L.chunks = fj_cdiv(L.n_ltasks > 0 ? L.n_ltasks : 1, FJ_BS_WAVES);
    if ((long long)L.tiles * L.chunks <= FJ_BS_TARGET_BLOCKS || L.width >= max_cols) break;
    L.width *= 2;
  }
  return L;
}

static bool fj_batched_lattice_form(const FjTaskDesc* d, int n, int precision) {
  if (precision != FJ_PREC_FP64) return false;
  for (int i = 0; i < n; ++i)
    if (d[i].n_tgt > 0 && (d[i].n_lat < 1 || d[i].n_lat > FJ_TASK_MAX_LAT)) return false;
  return true;
}

size_t fj_batched_ws_bytes(const FjTaskDesc* d, int n, int precision) {
  const bool lat = fj_batched_lattice_form(d, n, precision);
  size_t total = 0, recs = 0;
  for (int i = 0; i < n; ++i) {
    int tiles, chunks, chunk_len;
    if (lat) {
      const FjLatTiling L = fj_task_lattice_tiling(d[i]);
      tiles = L.tiles;
      chunks = L.chunks;
      recs += (size_t)L.n_recs;
    } else {
      fj_task_tiling(d[i].n_tgt, d[i].n_seg, tiles, chunks, chunk_len);
    }
    total += (size_t)chunks * 3 * (tiles > 0 ? tiles : 1) * 64;
  }
  return ((total * sizeof(double) + 255) / 256) * 256 + recs * FJ_LAT_REC * sizeof(double) + 256;
}

template <int KIND>
static void fj_batched_partial_launch(int precision, int grid, const FjTask* dtasks,
                                      const int* dcounts, const double* seg, double* part,
                                      const FjKernelParams& kp, hipStream_t st) {
  if (precision == FJ_PREC_MIXED)
    hipLaunchKernelGGL((k_batched_partial<KIND, true>), dim3(grid), dim3(FJ_BS_THREADS), 0, st,
                       dtasks, dcounts, seg, part, kp);
  else
    hipLaunchKernelGGL((k_batched_partial<KIND, false>), dim3(grid), dim3(FJ_BS_THREADS), 0, st,
                       dtasks, dcounts, seg, part, kp);
}

int fj_batched_launch(const FjTaskDesc* d, int n_tasks, const double* seg, int kernel,
                      int precision, const FjKernelParams& kp, void* task_table,
                      void* host_staging, void* workspace, hipStream_t st, bool launch) {
  if (n_tasks <= 0) return FJ_OK;
  if (!d || !task_table || !host_staging || !workspace) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (precision != FJ_PREC_FP64 && precision != FJ_PREC_MIXED) return FJ_ERR_ARG;
  // host_staging: fj_batched_task_bytes(n_tasks) bytes of host memory that
  // stays valid until the copy has run (the caller keeps it)
  const bool lat = fj_batched_lattice_form(d, n_tasks, precision);
  FjTask* h = (FjTask*)host_staging;
  int work = 0, part = 0, n_tgt_total = 0, recs = 0;
  for (int i = 0; i < n_tasks; ++i) {
    FjTask& T = h[i];
    if (d[i].n_tgt < 0 || d[i].n_seg < 0 || d[i].mode < FJ_OUT_STORE || d[i].mode > FJ_OUT_CONVECT)
      return FJ_ERR_ARG;
    if (d[i].n_tgt > 0 && (!d[i].tgt || !d[i].out || (d[i].n_seg > 0 && !seg))) return FJ_ERR_ARG;
    T.tgt = d[i].tgt;
    T.out = d[i].out;
    T.tgt_off = n_tgt_total;
    T.n_tgt = d[i].n_tgt;
    T.seg_off = d[i].seg_off;
    T.n_seg = d[i].n_seg;
    T.n_lat = 0;
    T.n_ltasks = 0;
    T.n_recs = 0;
    T.rec_off = recs;
    T.cb_width = 1;
    if (lat) {
      const FjLatTiling L = fj_task_lattice_tiling(d[i]);
      T.tiles = L.tiles;
      T.chunks = L.chunks;
      T.chunk_len = 0;
      T.n_lat = T.n_tgt > 0 ? d[i].n_lat : 0;
      for (int k = 0; k < FJ_TASK_MAX_LAT; ++k) {
        const bool on = k < T.n_lat;
        T.lat_off[k] = on ? d[i].lat_off[k] : 0;
        T.lat_m[k] = on ? d[i].lat_m[k] : 0;
        T.lat_n[k] = on ? d[i].lat_n[k] : 0;
        T.lat_cb[k] = on ? L.cb[k] : 1;
        T.lat_task[k] = on ? L.ltask[k] : L.n_ltasks;
        T.lat_rec[k] = recs + (on ? L.rec[k] : L.n_recs);
      }
      T.cb_width = L.width;
      T.n_ltasks = L.n_ltasks;
      T.n_recs = L.n_recs;
      recs += L.n_recs;
    } else {
      fj_task_tiling(T.n_tgt, T.n_seg, T.tiles, T.chunks, T.chunk_len);
    }
    T.work_off = work;
    T.part_off = part;
    T.ld = T.tiles * 64;
    T.mode = d[i].mode;
    T.dt = d[i].dt;
    T.time = d[i].time;
    T.flow = d[i].flow;
    work += T.tiles * T.chunks;
    part += T.chunks * 3 * T.ld;
    n_tgt_total += T.n_tgt;
  }
  int* counts = (int*)(h + n_tasks);
  counts[0] = n_tasks;
  counts[1] = work;
  counts[2] = n_tgt_total;
  counts[3] = recs;
  const int part_total = part;
  if (!launch) return FJ_OK;
  if (hipMemcpyAsync(task_table, h, fj_batched_task_bytes(n_tasks), hipMemcpyHostToDevice, st) !=
      hipSuccess)
    return FJ_ERR_LAUNCH;
  const FjTask* dtasks = (const FjTask*)task_table;
  const int* dcounts = (const int*)(dtasks + n_tasks);
  const int grid = work < 4096 ? work : 4096;
  if (work > 0 && lat) {
    double* part = (double*)workspace;
    double* tab = (double*)((char*)workspace + (((size_t)part_total * sizeof(double) + 255) / 256) * 256);
    const long long words = (long long)recs * FJ_LAT_REC;
    const int rgrid = (int)(fj_cdiv(words, 256) < 65536 ? fj_cdiv(words, 256) : 65536);
    hipLaunchKernelGGL(k_batched_lattice_records, dim3(rgrid), dim3(256), 0, st, dtasks, dcounts,
                       seg, tab, kernel);
    FJ_CHECK_LAUNCH();
    if (kernel == FJ_KERNEL_EPSILON)
      hipLaunchKernelGGL(k_batched_lattice_partial<FJ_KERNEL_EPSILON>, dim3(grid), dim3(FJ_BS_THREADS), 0, st, dtasks, dcounts, tab, part, kp);
    else if (kernel == FJ_KERNEL_CUTOFF)
      hipLaunchKernelGGL(k_batched_lattice_partial<FJ_KERNEL_CUTOFF>, dim3(grid), dim3(FJ_BS_THREADS), 0, st, dtasks, dcounts, tab, part, kp);
    else
      hipLaunchKernelGGL(k_batched_lattice_partial<FJ_KERNEL_PLAIN>, dim3(grid), dim3(FJ_BS_THREADS), 0, st, dtasks, dcounts, tab, part, kp);
    FJ_CHECK_LAUNCH();
  } else if (work > 0) {
    double* part = (double*)workspace;
    if (kernel == FJ_KERNEL_EPSILON)
      fj_batched_partial_launch<FJ_KERNEL_EPSILON>(precision, grid, dtasks, dcounts, seg, part, kp, st);
    else if (kernel == FJ_KERNEL_CUTOFF)
      fj_batched_partial_launch<FJ_KERNEL_CUTOFF>(precision, grid, dtasks, dcounts, seg, part, kp, st);
    else
      fj_batched_partial_launch<FJ_KERNEL_PLAIN>(precision, grid, dtasks, dcounts, seg, part, kp, st);
    FJ_CHECK_LAUNCH();
  }
  if (n_tgt_total > 0) {
    hipLaunchKernelGGL(k_batched_reduce, dim3(fj_cdiv(n_tgt_total, 64)), dim3(64 * FJ_BRED_WAVES),
                       0, st, dtasks, dcounts, (const double*)workspace);
    FJ_CHECK_LAUNCH();
  }
  return FJ_OK;
}

extern "C" size_t fj_batched_workspace_bytes(int n_tasks, const int* n_tgt, const int* n_seg) {
  size_t total = 0;
  for (int i = 0; i < n_tasks; ++i) {
    int tiles, chunks, chunk_len;
    fj_task_tiling(n_tgt[i], n_seg[i], tiles, chunks, chunk_len);
    total += (size_t)chunks * 3 * (tiles > 0 ? tiles : 1) * 64;
  }
  return total * sizeof(double) + 256;
}

extern "C" int fj_pair_sum_batched(int n_tasks, const int* tgt_off, const int* n_tgt,
                                   const int* seg_off, const int* n_seg, const int* out_off,
                                   const int* mode, const double* dt, const double* time,
                                   const fj_flowfield* flows, const double* seg,
                                   const double* targets, double* out, int kernel, double eps,
                                   double rcut, void* task_table, void* host_staging,
                                   void* workspace, void* stream) {
  if (n_tasks <= 0) return FJ_OK;
  if (!tgt_off || !n_tgt || !seg_off || !n_seg || !out_off || !mode || !task_table ||
      !host_staging || !workspace || !targets || !out)
    return FJ_ERR_ARG;
  // the descriptors are built in the tail of host_staging (after the task
  // records), which fj_batched_task_bytes sizes for both
  FjTaskDesc* d = (FjTaskDesc*)((char*)host_staging + fj_batched_task_bytes(n_tasks) -
                                fj_batched_desc_bytes(n_tasks));
  for (int i = 0; i < n_tasks; ++i) {
    if (i > 0 && tgt_off[i] < tgt_off[i - 1] + n_tgt[i - 1]) return FJ_ERR_ARG;  // sorted, disjoint
    d[i].tgt = targets + 3 * (size_t)tgt_off[i];
    d[i].out = out + 3 * (size_t)out_off[i];
    d[i].n_tgt = n_tgt[i];
    d[i].seg_off = seg_off[i];
    d[i].n_seg = n_seg[i];
    d[i].mode = mode[i];
    d[i].dt = dt ? dt[i] : 0.0;
    d[i].time = time ? time[i] : 0.0;
    d[i].flow = fj_flow_from_desc(flows ? &flows[i] : nullptr);
    d[i].n_lat = 0;
  }
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  return fj_batched_launch(d, n_tasks, seg, kernel, FJ_PREC_FP64, kp, task_table, host_staging,
                           workspace, (hipStream_t)stream);
}

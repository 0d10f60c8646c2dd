// C ABI implementation (see include/bemhip.h): context, mesh upload, workspace, launch sequences.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bemhip.h"
#include "bem_kernels.hpp"

// ---- error handling ----
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int bem_fail_message(int code, const std::string& msg) { return fail(code, msg); }
static int hipfail(hipError_t e, const char* what) {
  return fail(BEM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(call, what)                         \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return hipfail(e_, what); \
  } while (0)

#define BEM_TIMER_SLOTS 8
#define BEM_AUX_STREAMS 3

// Static scratch of bem_refine_noalloc: one copy per device, held until the process ends. Host code never uses
// the array's name as an address: each device's copy is resolved once, with that device current, by
// hipGetSymbolAddress in bem_ctx_create, and that address is what the kernels receive.
static __device__ __align__(256) unsigned char scratch[BEM_NOALLOC_SCRATCH_BYTES];
#define BEM_MAX_DEVICES 64
static void* g_scratch_ptr[BEM_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int resolve_scratch(int device) {
  if (device < 0 || device >= BEM_MAX_DEVICES) return fail(BEM_E_ARG, "device index above the library's limit");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return BEM_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hipfail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = p;
  return BEM_OK;
}

static void* scratch_ptr(int device) {
  if (device < 0 || device >= BEM_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

struct bem_ctx {
  int device;
  hipStream_t stream;
  bool own_stream;
  hipGraph_t graph;
  bool capturing;
  bool timing;
  hipEvent_t ev0, ev1;
  double timer_ms[BEM_TIMER_SLOTS];
  long long timer_count[BEM_TIMER_SLOTS];
  long long timer_regions[BEM_TIMER_SLOTS];
  // auxiliary streams (default priority) for batch fan-out: independent matrices of one batch progress on
  // separate streams so that one group's latency-bound panel phase overlaps another group's MFMA updates
  hipStream_t aux[BEM_AUX_STREAMS];
  hipEvent_t fork_ev, join_ev[BEM_AUX_STREAMS];
  int fanout;  // streams used for batch fan-out (1 .. 1 + BEM_AUX_STREAMS)
  hipEvent_t lu_ev[2];  // look-ahead LU: ready (context stream), done (look-ahead stream)
  hipEvent_t asm_ev[4];  // assembly: strip set s written (2 s, context stream), gathered (2 s + 1, aux[0])
  int lu_products;      // BEM_LU_FP32, BEM_LU_SPLIT_BF16 or BEM_LU_BF16: large updates of FP32 factorizations
  int sing_precision;   // BEM_SING_WORKING or BEM_SING_FP64: arithmetic of the Duffy sums of mixed batches
  int num_cus;          // compute units of the device
  std::vector<bem_mesh*> capture_meshes;  // meshes used by the sequence being captured
  bem_tuning tune;                        // launch-shape choices (bem_ctx_set_tuning)
  // bem_solve_status: device flags and their host copy
  int* status_dev = nullptr;
  int status_cap = 0;
};

// Run body(first, count, stream) on up to 1 + BEM_AUX_STREAMS contiguous sub-batches, forked from and joined
// back into the context stream with events (capturable into a graph).
template <typename F>
static hipError_t fan_out(bem_ctx* c, int batch, int min_per_stream, F&& body) {
  int S = c->fanout;
  while (S > 1 && batch < S * min_per_stream) --S;
  if (S == 1) return body(0, batch, c->stream);
  hipError_t e = hipEventRecord(c->fork_ev, c->stream);
  if (e != hipSuccess) return e;
  for (int i = 0; i < S - 1; ++i)
    if ((e = hipStreamWaitEvent(c->aux[i], c->fork_ev, 0)) != hipSuccess) return e;
  for (int i = 0; i < S; ++i) {
    const int b0 = (int)((long long)batch * i / S), b1 = (int)((long long)batch * (i + 1) / S);
    if ((e = body(b0, b1 - b0, i == 0 ? c->stream : c->aux[i - 1])) != hipSuccess) return e;
  }
  for (int i = 0; i < S - 1; ++i) {
    if ((e = hipEventRecord(c->join_ev[i], c->aux[i])) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream, c->join_ev[i], 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// Kernel timing on the context's stream (diagnostics): brackets one launch with events and waits for it.
struct ScopedTimer {
  bem_ctx* c;
  int slot;
  long long units;
  bool on;
  ScopedTimer(bem_ctx* c_, int slot_, long long units_) : c(c_), slot(slot_), units(units_) {
    on = c->timing && !c->capturing;
    if (on) (void)hipEventRecord(c->ev0, c->stream);
  }
  ~ScopedTimer() {
    if (!on) return;
    (void)hipEventRecord(c->ev1, c->stream);
    if (hipEventSynchronize(c->ev1) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
        c->timer_ms[slot] += ms;
        c->timer_count[slot] += units;
        c->timer_regions[slot] += 1;
      }
    }
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct bem_mesh {
  bem_ctx* ctx;
  int F, V, n_chunks, L, P, n_rules, n_src, M;
  int R;        // symmetry transforms (identity first): element records and field sources exist per transform
  int sgn[12];  // their sign vectors
  double m11[9], m10[3];
  int nq;  // points of the regular rule per face (6, 3 or 1)
  std::vector<DevBuf> owned;
  // device arrays
  ElemRec<float>* rec32;
  ElemRec<double>* rec64;
  SrcRec<float>* src32;
  SrcRec<double>* src64;
  float* srcw32;
  double* srcw64;
  int4* fv;
  int *order, *slot_count, *slot_vertex, *slot_pend, *slot_prev, *elem_pos, *chunk_start, *vertex_last, *vstar_ptr,
      *vstar_row;
  int4* vstar_ent;       // per star entry: strip row, first chunk of its row / column strip, the star's vertex
  uint32_t* slot_bytes;  // per chunk: 256 bytes of incidence starts, 256 bytes of incidence codes
  uint16_t *slot_start, *slot_inc;
  double *fvtx, *fnrm, *fcurl, *fa2, *rpts;
  int *sing_test, *sing_trial, *sing_rule, *sing_ptr, *rule_off, *rule_len, *src_faces, *sing_by_rule;
  double* roww;  // [V] orbit row weights or null
  int* pmask;    // [F] image-pair participation bits or null
  // workspace
  DevBuf strip, sing_lhs, sing_rhs, rhs_part, wsrc, partial, cols_part;
  int strip_nf = 0;
  int strip_sets = 1;  // 2: the frequency groups alternate between two strip sets
  // captured graphs whose nodes address this mesh's buffers; while any exists the workspace may not move
  int live_graphs = 0;
  bool closed = false;  // bem_mesh_destroy was called while graphs were live
  // the singular RHS blocks in sing_rhs belong to these wavenumbers (bem_assemble_rhs reuses them)
  const double* sing_k = nullptr;
  int sing_nf = 0;
};

// Executable graph handle: the instantiated graph and the meshes its nodes use.
struct bem_graph {
  hipGraphExec_t exec;
  std::vector<bem_mesh*> meshes;
};

static int dev_alloc(bem_mesh* m, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return BEM_OK;
  if (m->ctx->capturing) return fail(BEM_E_STATE, "device allocation requested during graph capture; call bem_mesh_reserve first");
  if (m->live_graphs > 0)
    return fail(BEM_E_STATE, "the mesh workspace must grow, but a captured graph still uses it; destroy the graph "
                             "first or reserve the largest shape before capturing");
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return BEM_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return fail(BEM_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  b.bytes = bytes;
  return BEM_OK;
}

// A mesh's tables go to the device as one allocation: each table is packed at a 256-byte aligned offset of a host
// image, copied in one transfer, and the table pointers set into the allocation. One hipMalloc and one hipMemcpy
// per mesh instead of one of each per table (37 tables: 4.7 ms per 6,000-triangle mesh, most of it per-call cost).
struct TableArena {
  struct Entry {
    void** dst;
    const void* src;
    size_t bytes, off;
  };
  std::vector<Entry> entries;
  size_t total = 0;
  template <typename U>
  void add(U** dst, const U* src, size_t count) {
    *dst = nullptr;
    if (count == 0) return;
    entries.push_back({reinterpret_cast<void**>(dst), src, count * sizeof(U), total});
    total += (count * sizeof(U) + 255) / 256 * 256;
  }
  int commit(bem_mesh* m) {
    if (total == 0) return BEM_OK;
    std::vector<unsigned char> image(total, 0);
    for (const Entry& e : entries)
      if (e.src) std::memcpy(image.data() + e.off, e.src, e.bytes);
    DevBuf b;
    hipError_t err = hipMalloc(&b.p, total);
    if (err != hipSuccess) return fail(BEM_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(err));
    b.bytes = total;
    m->owned.push_back(b);
    err = hipMemcpy(b.p, image.data(), total, hipMemcpyHostToDevice);
    if (err != hipSuccess) return hipfail(err, "hipMemcpy");
    for (const Entry& e : entries) *e.dst = static_cast<unsigned char*>(b.p) + e.off;
    return BEM_OK;
  }
};

extern "C" {

int bem_version(void) { return 10000; }

int bem_last_error(char* buf, int buflen) {
  if (!buf || buflen <= 0) return (int)g_err.size();
  std::snprintf(buf, (size_t)buflen, "%s", g_err.c_str());
  return (int)g_err.size();
}

int bem_ctx_create(int device, void* stream, bem_ctx** out) {
  if (!out) return fail(BEM_E_ARG, "out is NULL");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (device < 0 || device >= n) return fail(BEM_E_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  // the static scratch of bem_refine_noalloc: this device's copy, resolved with the device current
  {
    const int rc = resolve_scratch(device);
    if (rc) return rc;
  }
  bem_ctx* c = new bem_ctx();
  c->device = device;
  c->capturing = false;
  c->graph = nullptr;
  c->timing = false;
  c->fanout = 1 + BEM_AUX_STREAMS;
  c->lu_products = BEM_LU_FP32;
  c->sing_precision = BEM_SING_WORKING;
  c->tune = bem_tuning{0, BEM_TSW_MIN_BATCH_DEFAULT, 0, 0};
  c->num_cus = 0;
  if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      c->num_cus <= 0)
    c->num_cus = 256;
  for (int i = 0; i < BEM_TIMER_SLOTS; ++i) {
    c->timer_ms[i] = 0.0;
    c->timer_count[i] = 0;
    c->timer_regions[i] = 0;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return fail(BEM_E_HIP, "hipEventCreate failed");
  }
  for (int i = 0; i < BEM_AUX_STREAMS; ++i) {
    if (hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming) != hipSuccess) {
      delete c;
      return fail(BEM_E_HIP, "auxiliary stream creation failed");
    }
  }
  for (int i = 0; i < 2; ++i)
    if (hipEventCreateWithFlags(&c->lu_ev[i], hipEventDisableTiming) != hipSuccess) {
      delete c;
      return fail(BEM_E_HIP, "hipEventCreate failed");
    }
  for (int i = 0; i < 4; ++i)
    if (hipEventCreateWithFlags(&c->asm_ev[i], hipEventDisableTiming) != hipSuccess) {
      delete c;
      return fail(BEM_E_HIP, "hipEventCreate failed");
    }
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
    c->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return hipfail(e, "hipStreamCreate");
    }
    c->own_stream = true;
  }
  *out = c;
  return BEM_OK;
}

int bem_ctx_destroy(bem_ctx* c) {
  if (!c) return BEM_OK;
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  for (int i = 0; i < BEM_AUX_STREAMS; ++i) {
    (void)hipStreamDestroy(c->aux[i]);
    (void)hipEventDestroy(c->join_ev[i]);
  }
  for (int i = 0; i < 2; ++i) (void)hipEventDestroy(c->lu_ev[i]);
  for (int i = 0; i < 4; ++i) (void)hipEventDestroy(c->asm_ev[i]);
  (void)hipEventDestroy(c->fork_ev);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  if (c->status_dev) (void)hipFree(c->status_dev);
  delete c;
  return BEM_OK;
}

int bem_ctx_sync(bem_ctx* c) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  HIPCHK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return BEM_OK;
}

int bem_ctx_set_fanout(bem_ctx* c, int streams) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  if (streams < 1 || streams > 1 + BEM_AUX_STREAMS) return fail(BEM_E_ARG, "fan-out must be 1..4 streams");
  if (c->capturing) return fail(BEM_E_STATE, "cannot change the fan-out during graph capture");
  c->fanout = streams;
  return BEM_OK;
}

int bem_ctx_set_lu_products(bem_ctx* c, int mode) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  if (mode != BEM_LU_FP32 && mode != BEM_LU_SPLIT_BF16 && mode != BEM_LU_BF16)
    return fail(BEM_E_ARG, "unknown LU product mode");
  c->lu_products = mode;
  return BEM_OK;
}

int bem_ctx_set_singular_precision(bem_ctx* c, int mode) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  if (mode != BEM_SING_WORKING && mode != BEM_SING_FP64) return fail(BEM_E_ARG, "unknown singular precision mode");
  if (c->capturing) return fail(BEM_E_STATE, "cannot change the singular precision during graph capture");
  c->sing_precision = mode;
  return BEM_OK;
}

int bem_ctx_set_tuning(bem_ctx* c, const bem_tuning* t) {
  if (!c || !t) return fail(BEM_E_ARG, "NULL argument");
  if (t->split_tile < 0 || t->split_tile > 1) return fail(BEM_E_ARG, "split_tile must be 0 or 1");
  if (t->trsv_wg_min_batch < 0) return fail(BEM_E_ARG, "trsv_wg_min_batch must be >= 0");
  if (t->lookahead_wgs < -1) return fail(BEM_E_ARG, "lookahead_wgs must be >= -1");
  if (t->field_slice_budget < 0) return fail(BEM_E_ARG, "field_slice_budget must be >= 0");
  if (c->capturing) return fail(BEM_E_STATE, "cannot change the tuning during graph capture");
  c->tune = *t;
  return BEM_OK;
}

int bem_ctx_get_tuning(bem_ctx* c, bem_tuning* t) {
  if (!c || !t) return fail(BEM_E_ARG, "NULL argument");
  *t = c->tune;
  return BEM_OK;
}

int bem_ctx_timers(bem_ctx* c, int enable) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  c->timing = enable != 0;
  if (enable)
    for (int i = 0; i < BEM_TIMER_SLOTS; ++i) {
      c->timer_ms[i] = 0.0;
      c->timer_count[i] = 0;
      c->timer_regions[i] = 0;
    }
  return BEM_OK;
}

int bem_ctx_timer_read(bem_ctx* c, int slot, double* ms, long long* units, long long* regions) {
  if (!c || !ms || !units || slot < 0 || slot >= BEM_TIMER_SLOTS) return fail(BEM_E_ARG, "invalid timer query");
  *ms = c->timer_ms[slot];
  *units = c->timer_count[slot];
  if (regions) *regions = c->timer_regions[slot];
  return BEM_OK;
}

int bem_ctx_stream(bem_ctx* c, void** s) {
  if (!c || !s) return fail(BEM_E_ARG, "NULL argument");
  *s = c->stream;
  return BEM_OK;
}

}  // extern "C"

// Every index table of a mesh description is checked here, on the host, before anything is uploaded: the
// kernels index device memory with these values and do no bounds checks of their own. Returns an empty string
// when the description is consistent, else what is wrong.
static std::string desc_problem(const bem_mesh_desc* d) {
  const long long F = d->n_faces, V = d->n_vertices;
  if (F <= 0 || V <= 0) return "empty mesh";
  if (F > (1 << 24) || V > (1 << 24)) return "more than 2^24 faces or vertices";
  if (!d->faces || !d->face_vertices || !d->normals || !d->curls || !d->area2 || !d->qpoints)
    return "NULL geometry array";
  if (!d->order || !d->slot_count || !d->slot_start || !d->slot_inc || !d->slot_vertex || !d->slot_pend ||
      !d->slot_prev || !d->elem_pos || !d->chunk_start || !d->vertex_last || !d->vstar_ptr || !d->vstar_row)
    return "NULL slot table";
  if (d->n_chunks != (d->n_faces + 63) / 64) return "n_chunks must be ceil(F/64)";
  const int R = d->n_transforms <= 0 ? 1 : d->n_transforms;
  if (R != 1 && R != 2 && R != 4) return "n_transforms must be 1, 2 or 4";
  for (int i = 0; i < 3 * R; ++i)
    if (d->n_transforms > 0 && d->transform_signs[i] != 1 && d->transform_signs[i] != -1)
      return "transform signs must be +1 or -1";
  if (d->n_transforms > 0)
    for (int i = 0; i < 3; ++i)
      if (d->transform_signs[i] != 1) return "transform 0 must be the identity";
  for (long long i = 0; i < 3 * F; ++i)
    if (d->faces[i] < 0 || d->faces[i] >= V) return "face vertex id out of range";
  for (long long t = 0; t < F; ++t) {
    const int* f = d->faces + 3 * t;
    if (f[0] == f[1] || f[0] == f[2] || f[1] == f[2]) return "face with a repeated vertex";
    if (!(d->area2[t] > 0.0)) return "face with non-positive area";
  }
  // trial order: a permutation of the faces followed by -1 padding, inverted by elem_pos
  const long long nord = 64LL * d->n_chunks;
  for (long long i = 0; i < nord; ++i) {
    const int e = d->order[i];
    if (i < F ? (e < 0 || e >= F) : e != -1) return "order must be a permutation of the faces padded with -1";
  }
  for (long long t = 0; t < F; ++t) {
    if (d->elem_pos[t] < 0 || d->elem_pos[t] >= F || d->order[d->elem_pos[t]] != t) return "elem_pos does not invert order";
    if (d->chunk_start[t] < 0 || d->chunk_start[t] > d->elem_pos[t] / 64) return "chunk_start out of range";
  }
  for (long long v = 0; v < V; ++v)
    if (d->vertex_last[v] < -1 || d->vertex_last[v] >= d->n_chunks) return "vertex_last out of range";
  for (int ch = 0; ch < d->n_chunks; ++ch) {
    if (d->slot_count[ch] < 0 || d->slot_count[ch] > BEM_SLOT_INC) return "slot count out of range";
    // incidence ranges non-decreasing within the chunk's 192 entries; codes (lane << 2 | corner) of 64 lanes
    const uint16_t* st = d->slot_start + (size_t)ch * BEM_SLOT_STRIDE;
    if (st[0] != 0) return "slot incidence ranges must start at 0";
    for (int k = 0; k < d->slot_count[ch]; ++k)
      if (st[k] > st[k + 1] || st[k + 1] > BEM_SLOT_INC) return "slot incidence range out of order";
    for (int e = 0; e < st[d->slot_count[ch]]; ++e) {
      const int code = d->slot_inc[(size_t)ch * BEM_SLOT_INC + e];
      if ((code >> 2) >= 64 || (code & 3) == 3) return "slot incidence code out of range";
    }
    for (int k = 0; k < d->slot_count[ch]; ++k) {
      const int v = d->slot_vertex[(size_t)ch * BEM_SLOT_INC + k] >> 1;
      const int pd = d->slot_pend[(size_t)ch * BEM_SLOT_INC + k] & 0xFFFF;
      if (v < 0 || v >= V) return "slot vertex out of range";
      if (pd != BEM_PEND_NONE && pd >= BEM_PEND_SLOTS) return "LDS partial-sum slot out of range";
      const int pv = d->slot_prev[(size_t)ch * BEM_SLOT_INC + k];
      if (pv < -1 || pv >= ch) return "slot_prev out of range";
    }
  }
  // vertex -> strip-row CSR: every strip row 3t+l once, under the vertex of corner l of face t
  if (d->vstar_ptr[0] != 0 || d->vstar_ptr[V] != 3 * F) return "vstar_ptr must run from 0 to 3F";
  for (long long v = 0; v < V; ++v) {
    if (d->vstar_ptr[v + 1] < d->vstar_ptr[v]) return "vstar_ptr not monotone";
    for (int q = d->vstar_ptr[v]; q < d->vstar_ptr[v + 1]; ++q) {
      const int r = d->vstar_row[q];
      if (r < 0 || r >= 3 * F || d->faces[r] != v) return "vstar_row entry out of range or not at its vertex";
    }
  }
  // singular pairs grouped by test element, trial keys s + r F, Duffy rules within the rule table
  const long long P = d->n_pairs;
  if (P < 0 || d->n_rules < 0) return "negative pair or rule count";
  if (P > 0 && (!d->sing_test || !d->sing_trial || !d->sing_rule)) return "NULL singular pair table";
  if (!d->sing_ptr || d->sing_ptr[0] != 0 || d->sing_ptr[F] != P) return "sing_ptr must run from 0 to n_pairs";
  for (long long t = 0; t < F; ++t) {
    if (d->sing_ptr[t + 1] < d->sing_ptr[t]) return "sing_ptr not monotone";
    for (int q = d->sing_ptr[t]; q < d->sing_ptr[t + 1]; ++q)
      if (d->sing_test[q] != t) return "singular pairs must be grouped by test element";
  }
  for (long long i = 0; i < P; ++i) {
    if (d->sing_trial[i] < 0 || d->sing_trial[i] >= R * F) return "singular trial key out of range";
    if (d->sing_rule[i] < 0 || d->sing_rule[i] >= d->n_rules) return "singular rule index out of range";
  }
  if (d->n_rules > 0 && (!d->rule_off || !d->rule_len || !d->rule_pts)) return "NULL rule table";
  for (int r = 0; r < d->n_rules; ++r)
    if (d->rule_off[r] < 0 || d->rule_len[r] <= 0 || (long long)d->rule_off[r] + d->rule_len[r] > (1LL << 26))
      return "rule offset or length out of range";
  if (d->n_qp != 0 && d->n_qp != 6 && d->n_qp != 3 && d->n_qp != 1)
    return "n_qp must be 6 (order 4), 3 (order 2) or 1 (order 1)";
  if (d->n_src < 0 || (d->n_src > 0 && !d->src_faces)) return "invalid radiating-face list";
  for (int i = 0; i < d->n_src; ++i)
    if (d->src_faces[i] < 0 || d->src_faces[i] >= F) return "radiating face out of range";
  return std::string();
}

extern "C" {

int bem_mesh_validate(const bem_mesh_desc* d) {
  if (!d) return fail(BEM_E_ARG, "NULL argument");
  try {
    const std::string why = desc_problem(d);
    if (!why.empty()) return fail(BEM_E_ARG, "invalid mesh description: " + why);
  } catch (const std::exception& ex) {
    return fail(BEM_E_NOMEM, std::string("mesh validation: ") + ex.what());
  }
  return BEM_OK;
}

}  // extern "C"

static int mesh_create_impl(bem_ctx* c, const bem_mesh_desc* d, bem_mesh** out);

extern "C" int bem_mesh_create(bem_ctx* c, const bem_mesh_desc* d, bem_mesh** out) {
  if (!c || !d || !out) return fail(BEM_E_ARG, "NULL argument");
  *out = nullptr;
  int rc = bem_mesh_validate(d);
  if (rc) return rc;
  try {
    return mesh_create_impl(c, d, out);
  } catch (const std::exception& ex) {
    return fail(BEM_E_NOMEM, std::string("bem_mesh_create: ") + ex.what());
  }
}

static int mesh_create_impl(bem_ctx* c, const bem_mesh_desc* d, bem_mesh** out) {
  const int R = d->n_transforms <= 0 ? 1 : d->n_transforms;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  bem_mesh* m = new bem_mesh();
  m->ctx = c;
  const int F = d->n_faces, V = d->n_vertices;
  m->F = F;
  m->V = V;
  m->n_chunks = d->n_chunks;
  m->L = (V + 63) / 64 * 64;  // strip rows hold one column per vertex
  m->P = d->n_pairs;
  m->n_rules = d->n_rules;
  m->n_src = d->n_src;
  m->R = R;
  const int nq = d->n_qp ? d->n_qp : 6;
  m->nq = nq;
  m->M = nq * F * R;
  for (int i = 0; i < 12; ++i) m->sgn[i] = i < 3 * R && d->n_transforms > 0 ? d->transform_signs[i] : 1;
  std::memcpy(m->m11, d->m11, sizeof(m->m11));
  std::memcpy(m->m10, d->m10, sizeof(m->m10));
  // host-side records: transform r of element t at r*F + t (reflected rule points and normal, curls times
  // det * signs: BeatEngineCore.jl:341-356); field sources transform-major, nq per element
  std::vector<ElemRec<float>> r32((size_t)F * R);
  std::vector<ElemRec<double>> r64((size_t)F * R);
  std::vector<SrcRec<float>> s32(nq * (size_t)F * R);
  std::vector<SrcRec<double>> s64(nq * (size_t)F * R);
  std::vector<float> w32(nq * (size_t)F);
  std::vector<double> w64(nq * (size_t)F);
  std::vector<int4> fv(F);
  for (int t = 0; t < F; ++t) {
    for (int i = 0; i < 18; ++i) {
      r64[t].q[i] = i < 3 * nq ? d->qpoints[18 * (size_t)t + i] : 0.0;
      r32[t].q[i] = (float)r64[t].q[i];
    }
    for (int i = 0; i < 3; ++i) {
      r64[t].n[i] = d->normals[3 * (size_t)t + i];
      r32[t].n[i] = (float)r64[t].n[i];
    }
    for (int i = 0; i < 9; ++i) {
      r64[t].curl[i] = d->curls[9 * (size_t)t + i];
      r32[t].curl[i] = (float)r64[t].curl[i];
    }
    r64[t].a2 = d->area2[t];
    r32[t].a2 = (float)d->area2[t];
    r64[t].pad = 0;
    r32[t].pad = 0;
    fv[t] = make_int4(d->faces[3 * (size_t)t], d->faces[3 * (size_t)t + 1], d->faces[3 * (size_t)t + 2], 0);
    for (int q = 0; q < nq; ++q) {
      const size_t mm = nq * (size_t)t + q;
      for (int i = 0; i < 3; ++i) {
        s64[mm].y[i] = r64[t].q[3 * q + i];
        s64[mm].n[i] = r64[t].n[i];
        s32[mm].y[i] = (float)s64[mm].y[i];
        s32[mm].n[i] = (float)s64[mm].n[i];
      }
      w64[mm] = qrule_w(nq, q) * d->area2[t];
      w32[mm] = (float)w64[mm];
    }
  }
  for (int r = 1; r < R; ++r) {
    const int* sg = m->sgn + 3 * r;
    const double det = double(sg[0] * sg[1] * sg[2]);
    for (int t = 0; t < F; ++t) {
      ElemRec<double>& e = r64[(size_t)r * F + t];
      e = r64[t];
      for (int q = 0; q < nq; ++q)
        for (int i = 0; i < 3; ++i) e.q[3 * q + i] *= sg[i];
      for (int i = 0; i < 3; ++i) e.n[i] *= sg[i];
      for (int l = 0; l < 3; ++l)
        for (int i = 0; i < 3; ++i) e.curl[3 * l + i] *= det * sg[i];
      ElemRec<float>& ef = r32[(size_t)r * F + t];
      for (int i = 0; i < 18; ++i) ef.q[i] = (float)e.q[i];
      for (int i = 0; i < 3; ++i) ef.n[i] = (float)e.n[i];
      for (int i = 0; i < 9; ++i) ef.curl[i] = (float)e.curl[i];
      ef.a2 = (float)e.a2;
      ef.pad = 0;
      for (int q = 0; q < nq; ++q) {
        const size_t mm = (size_t)r * nq * F + nq * (size_t)t + q;
        for (int i = 0; i < 3; ++i) {
          s64[mm].y[i] = e.q[3 * q + i];
          s64[mm].n[i] = e.n[i];
          s32[mm].y[i] = (float)s64[mm].y[i];
          s32[mm].n[i] = (float)s64[mm].n[i];
        }
      }
    }
  }
  int rc = BEM_OK;
  size_t total_rule_pts = 0;
  for (int r = 0; r < d->n_rules; ++r) total_rule_pts = std::max(total_rule_pts, (size_t)(d->rule_off[r] + d->rule_len[r]));
  size_t vs_nnz = (size_t)d->vstar_ptr[V];
  // (the arena copies at commit: every host table it is given lives until then)
  TableArena arena;
  std::vector<uint32_t> sb;
  std::vector<int4> ent;
  std::vector<int> by_rule;
#define UP(dst, src, n) arena.add(&(dst), (src), (n))
  UP(m->rec32, r32.data(), r32.size());
  UP(m->rec64, r64.data(), r64.size());
  UP(m->src32, s32.data(), s32.size());
  UP(m->src64, s64.data(), s64.size());
  UP(m->srcw32, w32.data(), w32.size());
  UP(m->srcw64, w64.data(), w64.size());
  UP(m->fv, fv.data(), (size_t)F);
  UP(m->order, d->order, (size_t)d->n_chunks * 64);
  UP(m->slot_count, d->slot_count, (size_t)d->n_chunks);
  UP(m->slot_start, d->slot_start, (size_t)d->n_chunks * 196);
  UP(m->slot_inc, d->slot_inc, (size_t)d->n_chunks * 192);
  UP(m->slot_vertex, d->slot_vertex, (size_t)d->n_chunks * 192);
  UP(m->slot_pend, d->slot_pend, (size_t)d->n_chunks * 192);
  UP(m->slot_prev, d->slot_prev, (size_t)d->n_chunks * 192);
  {
    // per chunk 512 bytes: incidence starts of the slots (bytes 0..192) and incidence codes (bytes 256..447), the
    // paired kernel's LDS copy of the chunk's slot table
    sb.assign((size_t)d->n_chunks * 128, 0u);
    uint8_t* b = reinterpret_cast<uint8_t*>(sb.data());
    for (int ch = 0; ch < d->n_chunks; ++ch) {
      const uint16_t* st = d->slot_start + (size_t)ch * BEM_SLOT_STRIDE;
      for (int k = 0; k <= d->slot_count[ch]; ++k) b[(size_t)ch * 512 + k] = (uint8_t)st[k];
      for (int e = 0; e < st[d->slot_count[ch]]; ++e)
        b[(size_t)ch * 512 + 256 + e] = (uint8_t)d->slot_inc[(size_t)ch * BEM_SLOT_INC + e];
    }
    UP(m->slot_bytes, sb.data(), sb.size());
  }
  UP(m->elem_pos, d->elem_pos, (size_t)F);
  UP(m->chunk_start, d->chunk_start, (size_t)F);
  UP(m->vertex_last, d->vertex_last, (size_t)V);
  UP(m->vstar_ptr, d->vstar_ptr, (size_t)V + 1);
  UP(m->vstar_row, d->vstar_row, vs_nnz);
  {
    ent.resize(vs_nnz);
    for (int v = 0; v < V; ++v)
      for (int q = d->vstar_ptr[v]; q < d->vstar_ptr[v + 1]; ++q) {
        const int t = d->vstar_row[q] / 3;
        ent[q] = make_int4(d->vstar_row[q], d->chunk_start[t], d->elem_pos[t] / 64, v);
      }
    UP(m->vstar_ent, ent.data(), vs_nnz);
  }
  UP(m->fvtx, d->face_vertices, (size_t)F * 9);
  UP(m->fnrm, d->normals, (size_t)F * 3);
  UP(m->fcurl, d->curls, (size_t)F * 9);
  UP(m->fa2, d->area2, (size_t)F);
  UP(m->sing_test, d->sing_test, (size_t)d->n_pairs);
  UP(m->sing_trial, d->sing_trial, (size_t)d->n_pairs);
  UP(m->sing_rule, d->sing_rule, (size_t)d->n_pairs);
  UP(m->sing_ptr, d->sing_ptr, (size_t)F + 1);
  {
    // pairs sorted by Duffy rule (stable: a counting sort over the rules), so a wave of the singular kernel shares
    // one rule
    by_rule.resize(d->n_pairs);
    std::vector<int> start((size_t)d->n_rules + 1, 0);
    for (int i = 0; i < d->n_pairs; ++i) ++start[d->sing_rule[i] + 1];
    for (int r = 0; r < d->n_rules; ++r) start[r + 1] += start[r];
    for (int i = 0; i < d->n_pairs; ++i) by_rule[start[d->sing_rule[i]]++] = i;
    UP(m->sing_by_rule, by_rule.data(), (size_t)d->n_pairs);
  }
  UP(m->rule_off, d->rule_off, (size_t)d->n_rules);
  UP(m->rule_len, d->rule_len, (size_t)d->n_rules);
  UP(m->rpts, d->rule_pts, total_rule_pts * 5);
  UP(m->src_faces, d->src_faces, (size_t)d->n_src);
  m->roww = nullptr;
  m->pmask = nullptr;
  if (R > 1 && d->row_weights) UP(m->roww, d->row_weights, (size_t)V);
  if (R > 1 && d->plane_mask) UP(m->pmask, d->plane_mask, (size_t)F);
#undef UP
  if ((rc = arena.commit(m))) {
    bem_mesh_destroy(m);
    return rc;
  }
  *out = m;
  return BEM_OK;
}

static void mesh_free(bem_mesh* m) {
  (void)hipSetDevice(m->ctx->device);
  for (auto& b : m->owned)
    if (b.p) (void)hipFree(b.p);
  for (DevBuf* b : {&m->strip, &m->sing_lhs, &m->sing_rhs, &m->rhs_part, &m->wsrc, &m->partial, &m->cols_part})
    if (b->p) (void)hipFree(b->p);
  delete m;
}

// A mesh used by a captured graph is freed only when the last such graph is destroyed: the graph's kernel nodes
// hold its table and workspace addresses.
static void note_capture_use(bem_ctx* c, bem_mesh* m) {
  if (c->capturing && std::find(c->capture_meshes.begin(), c->capture_meshes.end(), m) == c->capture_meshes.end())
    c->capture_meshes.push_back(m);
}

extern "C" {

int bem_mesh_destroy(bem_mesh* m) {
  if (!m) return BEM_OK;
  if (m->live_graphs > 0) {
    m->closed = true;  // freed by bem_graph_destroy of the last graph that uses it
    return BEM_OK;
  }
  mesh_free(m);
  return BEM_OK;
}

static int field_slices(const bem_mesh* m, int n_obs, int nf) {
  const int obs_blocks = (n_obs + 63) / 64;
  const int budget = m->ctx->tune.field_slice_budget > 0 ? m->ctx->tune.field_slice_budget : 2048;
  int s = budget / std::max(1, obs_blocks * nf);
  s = std::max(1, std::min(s, (m->M + 255) / 256));
  return s;
}

static int reserve_impl(bem_mesh* m, int precision, int nf, int strip_nf, int n_obs, int n_ch) {
  const size_t cb = precision == BEM_PREC_MIXED ? 8 : 16;
  strip_nf = std::max(1, std::min(strip_nf, nf));
  int rc;
  // FP32 without images assembles pairs once: a row strip and a column strip per frequency; two strip sets of up
  // to 4 GiB each when there is more than one frequency group (a group's gather overlaps the next group's assembly)
  const size_t nstrips = (precision == BEM_PREC_MIXED && m->R == 1) ? 2 : 1;
  const size_t set_bytes = nstrips * strip_nf * 3 * m->F * (size_t)m->L * cb;
  // the bound is judged on strip rows of one column per vertex (as the host's footprint estimate counts them)
  const size_t set_bound = nstrips * strip_nf * 3 * m->F * (((size_t)m->V + 63) / 64 * 64) * cb;
  const int nsets = (nf > strip_nf && set_bound <= (4ull << 30)) ? 2 : 1;
  if ((rc = dev_alloc(m, m->strip, nsets * set_bytes))) return rc;
  m->strip_sets = nsets;
  m->strip_nf = strip_nf;
  if ((rc = dev_alloc(m, m->sing_lhs, (size_t)nf * std::max(1, m->P) * 9 * 16))) return rc;
  if ((rc = dev_alloc(m, m->sing_rhs, (size_t)nf * std::max(1, m->P) * 3 * 16))) return rc;
  if ((rc = dev_alloc(m, m->rhs_part, (size_t)nf * 3 * m->F * BEM_MAX_CHANNELS * 16))) return rc;
  if (n_obs > 0) {
    if ((rc = dev_alloc(m, m->wsrc, (size_t)nf * n_ch * m->nq * m->F * 2 * cb))) return rc;
    const int ns = field_slices(m, n_obs, nf);
    if ((rc = dev_alloc(m, m->partial, (size_t)nf * ns * n_ch * n_obs * 16))) return rc;
  }
  return BEM_OK;
}

int bem_mesh_reserve(bem_mesh* m, int precision, int nf, int strip_nf, int n_obs, int n_ch) {
  if (!m || nf <= 0 || n_ch < 1 || n_ch > BEM_MAX_CHANNELS) return fail(BEM_E_ARG, "invalid reserve arguments");
  // during capture nothing may be allocated (dev_alloc refuses): only the strip grouping is set
  if (!m->ctx->capturing) HIPCHK(hipSetDevice(m->ctx->device), "hipSetDevice");
  return reserve_impl(m, precision, nf, strip_nf, n_obs, n_ch);
}

int bem_mesh_device_bytes(bem_mesh* m, long long* bytes) {
  if (!m || !bytes) return fail(BEM_E_ARG, "NULL argument");
  size_t s = 0;
  for (auto& b : m->owned) s += b.bytes;
  for (DevBuf* b : {&m->strip, &m->sing_lhs, &m->sing_rhs, &m->rhs_part, &m->wsrc, &m->partial, &m->cols_part})
    s += b->bytes;
  *bytes = (long long)s;
  return BEM_OK;
}

}  // extern "C"

template <typename T>
static int assemble_impl(bem_ctx* c, bem_mesh* m, int nf, const double* k, const void* qunit, int n_ch,
                         const void* rhs_scale, void* A, long long lda, long long strideA, void* A0, void* b,
                         long long ldb, long long strideB, double* rowmax, int precision) {
  int rc;
  if ((rc = reserve_impl(m, precision, nf, m->strip_nf > 0 ? m->strip_nf : std::min(nf, 4), 0, n_ch))) return rc;
  note_capture_use(c, m);
  m->sing_k = k;
  m->sing_nf = nf;
  hipStream_t st = c->stream;
  // singular blocks for all frequencies
  SingParams sp;
  sp.fvtx = m->fvtx;
  sp.fnrm = m->fnrm;
  sp.fcurl = m->fcurl;
  sp.fa2 = m->fa2;
  sp.test = m->sing_test;
  sp.trial = m->sing_trial;
  sp.rule = m->sing_rule;
  sp.by_rule = m->sing_by_rule;
  sp.rule_off = m->rule_off;
  sp.rule_len = m->rule_len;
  sp.rpts = m->rpts;
  sp.k = k;
  sp.lhs = static_cast<cplx<double>*>(m->sing_lhs.p);
  sp.rhs = static_cast<cplx<double>*>(m->sing_rhs.p);
  sp.P = m->P;
  std::memcpy(sp.m11, m->m11, sizeof(sp.m11));
  std::memcpy(sp.m10, m->m10, sizeof(sp.m10));
  sp.F = m->F;
  std::memcpy(sp.sgn, m->sgn, sizeof(sp.sgn));
  {
    ScopedTimer tm(c, BEM_TIMER_SINGULAR, (long long)nf);
    HIPCHK(launch_singular(sp, nf, precision == BEM_PREC_MIXED && c->sing_precision == BEM_SING_WORKING, st),
           "k_singular launch");
  }
  // rhs partial sums
  RhsParams<T> rp;
  rp.rec = precision == BEM_PREC_MIXED ? (const ElemRec<T>*)m->rec32 : (const ElemRec<T>*)m->rec64;
  rp.fv = m->fv;
  rp.src = m->src_faces;
  rp.qunit = static_cast<const cplx<double>*>(qunit);
  rp.sing_ptr = m->sing_ptr;
  rp.sing_trial = m->sing_trial;
  rp.sing_rhs = static_cast<const cplx<double>*>(m->sing_rhs.p);
  rp.k = k;
  rp.out = static_cast<cplx<double>*>(m->rhs_part.p);
  rp.F = m->F;
  rp.P = m->P;
  rp.n_src = m->n_src;
  rp.n_ch = n_ch;
  rp.R = m->R;
  rp.pmask = m->pmask;
  rp.nq = m->nq;
  {
    ScopedTimer tm(c, BEM_TIMER_RHS, (long long)nf);
    HIPCHK(launch_rhs<T>(rp, nf, st), "k_rhs_part launch");
  }
  // regular strips + gather, in strip-sized frequency groups. With more than one group the groups alternate
  // between two strip sets and each group's gather runs on an auxiliary stream, beside the next group's
  // assembly (the VALU-bound strips and the HBM-bound gather share the device); a strip set is reused once its
  // gather is done. Forks from and joins into the context stream only (capturable). Timed runs keep one stream.
  const long long fstride = 3LL * m->F * m->L;
  const size_t set_stride = (size_t)((sizeof(T) == 4 && m->R == 1) ? 2 : 1) * m->strip_nf * fstride;
  const int ngroups = (nf + m->strip_nf - 1) / m->strip_nf;
  const bool overlap = ngroups > 1 && m->strip_sets == 2 && !c->timing && c->fanout >= 2;
  for (int f0 = 0, gi = 0; f0 < nf; f0 += m->strip_nf, ++gi) {
    const int g = std::min(m->strip_nf, nf - f0);
    const int set = m->strip_sets == 2 ? (gi & 1) : 0;
    if (overlap && gi >= 2) HIPCHK(hipStreamWaitEvent(st, c->asm_ev[2 * set + 1], 0), "strip set reuse");
    RegParams<T> p;
    p.rec = rp.rec;
    p.fv = m->fv;
    p.order = m->order;
    p.slot_count = m->slot_count;
    p.slot_start = m->slot_start;
    p.slot_inc = m->slot_inc;
    p.slot_bytes = m->slot_bytes;
    p.slot_vertex = m->slot_vertex;
    p.slot_pend = m->slot_pend;
    p.slot_prev = m->slot_prev;
    p.elem_pos = m->elem_pos;
    p.chunk_start = m->chunk_start;
    p.sing_ptr = m->sing_ptr;
    p.sing_trial = m->sing_trial;
    p.sing_lhs = static_cast<const cplx<double>*>(m->sing_lhs.p) + (size_t)f0 * m->P * 9;
    p.k = k + f0;
    p.strip = static_cast<cplx<T>*>(m->strip.p) + set * set_stride;
    // paired assembly (FP32, no images, order-4 rule): the column strip follows the row strips of the group
    p.strip2 = (sizeof(T) == 4 && m->R == 1 && m->nq == 6) ? p.strip + (size_t)m->strip_nf * fstride : nullptr;
    p.nq = m->nq;
    p.strip_fstride = fstride;
    p.L = m->L;
    p.F = m->F;
    p.n_chunks = m->n_chunks;
    p.P = m->P;
    p.R = m->R;
    p.pmask = m->pmask;
    {
      ScopedTimer tm(c, BEM_TIMER_REGULAR, (long long)g);
      HIPCHK(launch_regular<T>(p, g, st), "k_regular_lhs launch");
    }
    GatherParams<T> gp;
    gp.strip = p.strip;
    gp.strip2 = p.strip2;
    gp.elem_pos = m->elem_pos;
    gp.chunk_start = m->chunk_start;
    gp.vertex_last = m->vertex_last;
    gp.vstar_ent = m->vstar_ent;
    gp.strip_fstride = fstride;
    gp.L = m->L;
    gp.vstar_ptr = m->vstar_ptr;
    gp.vstar_row = m->vstar_row;
    gp.rhs_part = static_cast<const cplx<double>*>(m->rhs_part.p) + (size_t)f0 * 3 * m->F * BEM_MAX_CHANNELS;
    gp.rhs_scale = static_cast<const cplx<double>*>(rhs_scale) + f0;
    gp.A = static_cast<cplx<T>*>(A) + (size_t)f0 * strideA;
    gp.A0 = A0 ? static_cast<cplx<T>*>(A0) + (size_t)f0 * strideA : nullptr;
    gp.lda = lda;
    gp.strideA = strideA;
    gp.b = static_cast<cplx<double>*>(b) + (size_t)f0 * strideB;
    gp.ldb = ldb;
    gp.strideB = strideB;
    gp.rowmax = rowmax + (size_t)f0 * m->V;
    gp.N = m->V;
    gp.F = m->F;
    gp.n_ch = n_ch;
    gp.roww = m->roww;
    if (overlap) {
      HIPCHK(hipEventRecord(c->asm_ev[2 * set], st), "strip set written");
      HIPCHK(hipStreamWaitEvent(c->aux[0], c->asm_ev[2 * set], 0), "gather fork");
      HIPCHK(launch_gather<T>(gp, g, c->aux[0]), "k_gather launch");
      HIPCHK(hipEventRecord(c->asm_ev[2 * set + 1], c->aux[0]), "strip set gathered");
    } else {
      ScopedTimer tm(c, BEM_TIMER_GATHER, (long long)g);
      HIPCHK(launch_gather<T>(gp, g, st), "k_gather launch");
    }
  }
  if (overlap) {  // join: the last group's gather (the one before it finished before that group's assembly)
    const int last = (ngroups - 1) & 1;
    HIPCHK(hipStreamWaitEvent(st, c->asm_ev[2 * last + 1], 0), "gather join");
    if (ngroups >= 2) HIPCHK(hipStreamWaitEvent(st, c->asm_ev[2 * (last ^ 1) + 1], 0), "gather join");
  }
  return BEM_OK;
}

extern "C" {

int bem_assemble(bem_ctx* c, bem_mesh* m, int precision, int nf, const double* k, const void* qunit, int n_ch,
                 const void* rhs_scale, void* A, long long lda, long long strideA, void* A0, void* b,
                 long long ldb, long long strideB, double* rowmax) {
  if (!c || !m || !k || !qunit || !rhs_scale || !A || !b || !rowmax) return fail(BEM_E_ARG, "NULL argument");
  if (nf <= 0 || n_ch < 1 || n_ch > BEM_MAX_CHANNELS) return fail(BEM_E_ARG, "invalid nf or n_ch");
  if (lda < m->V || strideA < lda * m->V || ldb < m->V || strideB < ldb * n_ch)
    return fail(BEM_E_ARG, "leading dimension or stride too small");
  if (precision == BEM_PREC_MIXED)
    return assemble_impl<float>(c, m, nf, k, qunit, n_ch, rhs_scale, A, lda, strideA, A0, b, ldb, strideB, rowmax,
                                precision);
  if (precision == BEM_PREC_FP64)
    return assemble_impl<double>(c, m, nf, k, qunit, n_ch, rhs_scale, A, lda, strideA, A0, b, ldb, strideB,
                                 rowmax, precision);
  return fail(BEM_E_ARG, "unknown precision");
}

int bem_assemble_rhs(bem_ctx* c, bem_mesh* m, int precision, int nf, const double* k, const void* qunit, int n_ch,
                     const void* rhs_scale, void* b, long long ldb, long long strideB) {
  if (!c || !m || !k || !qunit || !rhs_scale || !b) return fail(BEM_E_ARG, "NULL argument");
  if (nf <= 0 || n_ch < 1 || n_ch > BEM_MAX_CHANNELS) return fail(BEM_E_ARG, "invalid nf or n_ch");
  if (ldb < m->V || strideB < ldb * n_ch) return fail(BEM_E_ARG, "leading dimension or stride too small");
  if (precision != BEM_PREC_MIXED && precision != BEM_PREC_FP64) return fail(BEM_E_ARG, "unknown precision");
  if (m->sing_k != k || m->sing_nf < nf)
    return fail(BEM_E_STATE, "bem_assemble_rhs reuses the singular blocks of the last bem_assemble on this mesh, "
                             "which was called with other wavenumbers");
  note_capture_use(c, m);
  hipStream_t st = c->stream;
  hipError_t e;
  auto run = [&](auto tag) -> hipError_t {
    using T = decltype(tag);
    RhsParams<T> rp;
    rp.rec = sizeof(T) == 4 ? (const ElemRec<T>*)m->rec32 : (const ElemRec<T>*)m->rec64;
    rp.fv = m->fv;
    rp.src = m->src_faces;
    rp.qunit = static_cast<const cplx<double>*>(qunit);
    rp.sing_ptr = m->sing_ptr;
    rp.sing_trial = m->sing_trial;
    rp.sing_rhs = static_cast<const cplx<double>*>(m->sing_rhs.p);
    rp.k = k;
    rp.out = static_cast<cplx<double>*>(m->rhs_part.p);
    rp.F = m->F;
    rp.P = m->P;
    rp.n_src = m->n_src;
    rp.n_ch = n_ch;
    rp.R = m->R;
    rp.pmask = m->pmask;
    rp.nq = m->nq;
    hipError_t e2;
    {
      ScopedTimer tm(c, BEM_TIMER_RHS, (long long)nf);
      if ((e2 = launch_rhs<T>(rp, nf, st)) != hipSuccess) return e2;
    }
    GatherParams<T> gp{};
    gp.vstar_ptr = m->vstar_ptr;
    gp.vstar_row = m->vstar_row;
    gp.rhs_part = rp.out;
    gp.rhs_scale = static_cast<const cplx<double>*>(rhs_scale);
    gp.b = static_cast<cplx<double>*>(b);
    gp.ldb = ldb;
    gp.strideB = strideB;
    gp.rowmax = nullptr;
    gp.N = m->V;
    gp.F = m->F;
    gp.n_ch = n_ch;
    gp.roww = m->roww;
    ScopedTimer tm(c, BEM_TIMER_GATHER, (long long)nf);
    return launch_gather_rhs<T>(gp, nf, st);
  };
  e = precision == BEM_PREC_MIXED ? run(float()) : run(double());
  if (e != hipSuccess) return hipfail(e, "assemble_rhs launch");
  return BEM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// Coupled FEM-BEM system blocks
// ------------------------------------------------------------------------------------------------

struct bem_colset {
  bem_mesh* mesh;
  int n_src = 0, n_cols = 0;
  long long nnz = 0;
  int* src = nullptr;              // [n_src] face ids
  int* col_ptr = nullptr;          // [n_cols + 1]
  int* col_idx = nullptr;          // [nnz] positions in src
  cplx<double>* col_w = nullptr;   // [nnz]
};

struct bem_sparse {
  int n_rows = 0, n_cols = 0;
  long long nnz = 0;
  int* rows = nullptr;
  int* cols = nullptr;
  cplx<double>* vals = nullptr;
};

template <typename U>
static int raw_upload(U** dst, const U* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return BEM_OK;
  hipError_t e = hipMalloc((void**)dst, count * sizeof(U));
  if (e != hipSuccess) return fail(BEM_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  e = hipMemcpy(*dst, src, count * sizeof(U), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hipfail(e, "upload");
  return BEM_OK;
}

static void colset_free(bem_colset* cs) {
  if (!cs) return;
  for (void* p : {(void*)cs->src, (void*)cs->col_ptr, (void*)cs->col_idx, (void*)cs->col_w})
    if (p) (void)hipFree(p);
  delete cs;
}

static void sparse_free(bem_sparse* sp) {
  if (!sp) return;
  for (void* p : {(void*)sp->rows, (void*)sp->cols, (void*)sp->vals})
    if (p) (void)hipFree(p);
  delete sp;
}

extern "C" {

int bem_colset_create(bem_ctx* c, bem_mesh* m, int n_src, const int* src, int n_cols, const int* col_ptr,
                      const int* col_idx, const double* col_w, bem_colset** out) {
  if (!c || !m || !out || (n_src > 0 && !src) || n_cols < 0 || n_src < 0 || !col_ptr)
    return fail(BEM_E_ARG, "NULL argument or negative size");
  *out = nullptr;
  if (col_ptr[0] != 0) return fail(BEM_E_ARG, "col_ptr must start at 0");
  for (int j = 0; j < n_cols; ++j)
    if (col_ptr[j + 1] < col_ptr[j]) return fail(BEM_E_ARG, "col_ptr not monotone");
  const long long nnz = col_ptr[n_cols];
  if (nnz > 0 && (!col_idx || !col_w)) return fail(BEM_E_ARG, "NULL column entries");
  for (int j = 0; j < n_src; ++j)
    if (src[j] < 0 || src[j] >= m->F) return fail(BEM_E_ARG, "column source face out of range");
  for (long long e = 0; e < nnz; ++e)
    if (col_idx[e] < 0 || col_idx[e] >= n_src) return fail(BEM_E_ARG, "column entry outside the source list");
  if (c->capturing) return fail(BEM_E_STATE, "bem_colset_create allocates; not during graph capture");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  bem_colset* cs = new bem_colset();
  cs->mesh = m;
  cs->n_src = n_src;
  cs->n_cols = n_cols;
  cs->nnz = nnz;
  int rc;
  if ((rc = raw_upload(&cs->src, src, (size_t)n_src)) || (rc = raw_upload(&cs->col_ptr, col_ptr, (size_t)n_cols + 1)) ||
      (rc = raw_upload(&cs->col_idx, col_idx, (size_t)nnz)) ||
      (rc = raw_upload(&cs->col_w, reinterpret_cast<const cplx<double>*>(col_w), (size_t)nnz))) {
    colset_free(cs);
    return rc;
  }
  *out = cs;
  return BEM_OK;
}

int bem_colset_destroy(bem_colset* cs) {
  colset_free(cs);
  return BEM_OK;
}

int bem_rhs_columns(bem_ctx* c, bem_mesh* m, const bem_colset* cs, int precision, int nf, const double* k,
                    const void* alpha, void* out, long long ldo, long long strideO) {
  if (!c || !m || !cs || !k || !alpha || !out) return fail(BEM_E_ARG, "NULL argument");
  if (cs->mesh != m) return fail(BEM_E_ARG, "the column set belongs to another mesh");
  if (nf <= 0) return fail(BEM_E_ARG, "invalid nf");
  if (precision != BEM_PREC_MIXED && precision != BEM_PREC_FP64) return fail(BEM_E_ARG, "unknown precision");
  if (ldo < cs->n_cols || strideO < ldo * m->V) return fail(BEM_E_ARG, "leading dimension or stride too small");
  if (m->sing_k != k || m->sing_nf < nf)
    return fail(BEM_E_STATE, "bem_rhs_columns reuses the singular blocks of the last bem_assemble on this mesh, "
                             "which was called with other wavenumbers");
  if (cs->n_cols == 0) return BEM_OK;
  int rc;
  if ((rc = dev_alloc(m, m->cols_part, std::max<size_t>(1, (size_t)nf * 3 * m->F * cs->n_src * 16)))) return rc;
  note_capture_use(c, m);
  hipStream_t st = c->stream;
  auto run = [&](auto tag) -> hipError_t {
    using T = decltype(tag);
    hipError_t e2;
    if (cs->n_src > 0) {
      RhsParams<T> rp{};
      rp.rec = sizeof(T) == 4 ? (const ElemRec<T>*)m->rec32 : (const ElemRec<T>*)m->rec64;
      rp.fv = m->fv;
      rp.src = cs->src;
      rp.sing_ptr = m->sing_ptr;
      rp.sing_trial = m->sing_trial;
      rp.sing_rhs = static_cast<const cplx<double>*>(m->sing_rhs.p);
      rp.k = k;
      rp.out = static_cast<cplx<double>*>(m->cols_part.p);
      rp.F = m->F;
      rp.P = m->P;
      rp.n_src = cs->n_src;
      rp.n_ch = 0;
      rp.R = m->R;
      rp.pmask = m->pmask;
      rp.nq = m->nq;
      ScopedTimer tm(c, BEM_TIMER_RHS, (long long)nf);
      if ((e2 = launch_rhs_cols<T>(rp, nf, st)) != hipSuccess) return e2;
    }
    ScopedTimer tm(c, BEM_TIMER_GATHER, (long long)nf);
    return cols_gather<T>(static_cast<const cplx<double>*>(m->cols_part.p), m->F, cs->n_src, m->vstar_ptr,
                          m->vstar_row, m->roww, m->V, cs->n_cols, cs->col_ptr, cs->col_idx, cs->col_w,
                          static_cast<const cplx<double>*>(alpha), static_cast<cplx<T>*>(out), ldo, strideO, nf, st);
  };
  hipError_t e = precision == BEM_PREC_MIXED ? run(float()) : run(double());
  if (e != hipSuccess) return hipfail(e, "rhs_columns launch");
  return BEM_OK;
}

int bem_sparse_create(bem_ctx* c, int n_rows, int n_cols, long long nnz, const int* rows, const int* cols,
                      const double* vals, bem_sparse** out) {
  if (!c || !out || n_rows < 0 || n_cols < 0 || nnz < 0 || (nnz > 0 && (!rows || !cols || !vals)))
    return fail(BEM_E_ARG, "NULL argument or negative size");
  *out = nullptr;
  std::vector<long long> keys((size_t)nnz);
  for (long long e = 0; e < nnz; ++e) {
    if (rows[e] < 0 || rows[e] >= n_rows || cols[e] < 0 || cols[e] >= n_cols)
      return fail(BEM_E_ARG, "sparse entry outside the block");
    keys[(size_t)e] = (long long)rows[e] * n_cols + cols[e];
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
    return fail(BEM_E_ARG, "sparse entries must be unique (sum duplicates first)");
  if (c->capturing) return fail(BEM_E_STATE, "bem_sparse_create allocates; not during graph capture");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  bem_sparse* sp = new bem_sparse();
  sp->n_rows = n_rows;
  sp->n_cols = n_cols;
  sp->nnz = nnz;
  int rc;
  if ((rc = raw_upload(&sp->rows, rows, (size_t)nnz)) || (rc = raw_upload(&sp->cols, cols, (size_t)nnz)) ||
      (rc = raw_upload(&sp->vals, reinterpret_cast<const cplx<double>*>(vals), (size_t)nnz))) {
    sparse_free(sp);
    return rc;
  }
  *out = sp;
  return BEM_OK;
}

int bem_sparse_destroy(bem_sparse* sp) {
  sparse_free(sp);
  return BEM_OK;
}

int bem_sparse_scatter(bem_ctx* c, const bem_sparse* sp, int precision, double alpha_re, double alpha_im, int add,
                       void* out, long long ldo) {
  if (!c || !sp || !out) return fail(BEM_E_ARG, "NULL argument");
  if (ldo < sp->n_cols) return fail(BEM_E_ARG, "leading dimension smaller than the block");
  hipError_t e;
  const cplx<double> a = mk<double>(alpha_re, alpha_im);
  if (precision == BEM_PREC_MIXED)
    e = scatter_coo<float>(sp->nnz, sp->rows, sp->cols, sp->vals, a, add, static_cast<cplx<float>*>(out), ldo,
                           c->stream);
  else if (precision == BEM_PREC_FP64)
    e = scatter_coo<double>(sp->nnz, sp->rows, sp->cols, sp->vals, a, add, static_cast<cplx<double>*>(out), ldo,
                            c->stream);
  else
    return fail(BEM_E_ARG, "unknown precision");
  if (e != hipSuccess) return hipfail(e, "sparse scatter launch");
  return BEM_OK;
}

int bem_dense_block(bem_ctx* c, int precision, int m, int n, const void* src, long long lds, double alpha_re,
                    double alpha_im, int add, void* out, long long ldo) {
  if (!c || !out || m < 0 || n < 0) return fail(BEM_E_ARG, "NULL argument or negative size");
  if (ldo < n || (src && lds < n)) return fail(BEM_E_ARG, "leading dimension smaller than the block");
  hipError_t e;
  const cplx<double> a = mk<double>(alpha_re, alpha_im);
  const cplx<double>* s = static_cast<const cplx<double>*>(src);
  if (precision == BEM_PREC_MIXED)
    e = dense_block<float>(m, n, s, lds, a, add, static_cast<cplx<float>*>(out), ldo, c->stream);
  else if (precision == BEM_PREC_FP64)
    e = dense_block<double>(m, n, s, lds, a, add, static_cast<cplx<double>*>(out), ldo, c->stream);
  else
    return fail(BEM_E_ARG, "unknown precision");
  if (e != hipSuccess) return hipfail(e, "dense block launch");
  return BEM_OK;
}

int bem_row_equilibrate(bem_ctx* c, int precision, int n, int ncols, void* A, long long lda, double* scale) {
  if (!c || !A || !scale) return fail(BEM_E_ARG, "NULL argument");
  if (n < 0 || ncols < 0 || lda < ncols) return fail(BEM_E_ARG, "invalid sizes");
  hipError_t e;
  if (precision == BEM_PREC_MIXED)
    e = row_equilibrate<float>(static_cast<cplx<float>*>(A), lda, n, ncols, scale, c->stream);
  else if (precision == BEM_PREC_FP64)
    e = row_equilibrate<double>(static_cast<cplx<double>*>(A), lda, n, ncols, scale, c->stream);
  else
    return fail(BEM_E_ARG, "unknown precision");
  if (e != hipSuccess) return hipfail(e, "row equilibration launch");
  return BEM_OK;
}

int bem_neumann(bem_ctx* c, int nf, int n_ch, int n_faces, const void* rhs_scale, const void* qunit, void* q) {
  if (!c || !rhs_scale || !qunit || !q) return fail(BEM_E_ARG, "NULL argument");
  if (nf <= 0 || n_ch <= 0 || n_faces <= 0) return fail(BEM_E_ARG, "invalid sizes");
  hipError_t e = neumann_batch(static_cast<const cplx<double>*>(rhs_scale), static_cast<const cplx<double>*>(qunit),
                               static_cast<cplx<double>*>(q), n_ch, n_faces, nf, c->stream);
  if (e != hipSuccess) return hipfail(e, "neumann launch");
  return BEM_OK;
}

int bem_solve_status(bem_ctx* c, int n, int batch, const int* info, const void* x, long long ldx, long long strideX,
                     int nrhs, const double* rel, double tol, const double* growth, double growth_max, int* flags) {
  if (!c || !info || !x || !flags) return fail(BEM_E_ARG, "NULL argument");
  if (n <= 0 || batch <= 0 || nrhs < 1 || ldx < n || strideX < ldx * nrhs) return fail(BEM_E_ARG, "invalid sizes");
  if (c->capturing) return fail(BEM_E_STATE, "bem_solve_status synchronizes and cannot be captured");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  if (c->status_cap < batch) {
    if (c->status_dev) (void)hipFree(c->status_dev);
    c->status_dev = nullptr;
    c->status_cap = 0;
    hipError_t e = hipMalloc(&c->status_dev, (size_t)batch * sizeof(int));
    if (e != hipSuccess) return fail(BEM_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->status_cap = batch;
  }
  HIPCHK(solve_status_batch(info, static_cast<const cplx<double>*>(x), ldx, strideX, n, nrhs, rel, tol, growth,
                            growth_max, c->status_dev, batch, c->stream),
         "solve_status launch");
  HIPCHK(hipMemcpyAsync(flags, c->status_dev, (size_t)batch * sizeof(int), hipMemcpyDeviceToHost, c->stream),
         "hipMemcpyAsync");
  HIPCHK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  int any = 0;
  for (int i = 0; i < batch; ++i) any |= flags[i];
  if (any & 3) {
    int first = 0;
    while (!(flags[first] & 3)) ++first;
    return fail(BEM_E_SINGULAR, "matrix " + std::to_string(first) +
                                    ((flags[first] & 1) ? ": LU breakdown (zero pivot)" : ": non-finite solution"));
  }
  if (any & 12) {
    int first = 0;
    while (!(flags[first] & 12)) ++first;
    return fail(BEM_E_ACCURACY, "matrix " + std::to_string(first) +
                                    ((flags[first] & 4) ? ": relative residual above the tolerance"
                                                        : ": pivot growth above the limit"));
  }
  return BEM_OK;
}

int bem_lu_workspace_bytes(int precision, int n, int batch, long long* bytes) {
  if (!bytes || n <= 0 || batch <= 0) return fail(BEM_E_ARG, "invalid arguments");
  const long long esz = precision == BEM_PREC_MIXED ? 8 : precision == BEM_PREC_FP64 ? 16 : 0;
  if (!esz) return fail(BEM_E_ARG, "unknown precision");
  *bytes = lu_workspace_elems(n) * esz * batch;
  return BEM_OK;
}

int bem_lu_factor(bem_ctx* c, int precision, int n, int batch, void* A, long long lda, long long strideA,
                  int* ipiv, int* perm, long long strideP, int* info, void* work) {
  if (!c || !A || !ipiv || !perm || !info || !work) return fail(BEM_E_ARG, "NULL argument");
  if (n <= 0 || batch <= 0 || lda < n || strideA < lda * n || strideP < n) return fail(BEM_E_ARG, "invalid sizes");
  if (n > BEM_LU_MAX_N) return fail(BEM_E_ARG, "N above 24576 is not supported by the register panel");
  if (precision != BEM_PREC_MIXED && precision != BEM_PREC_FP64) return fail(BEM_E_ARG, "unknown precision");
  ScopedTimer tm(c, BEM_TIMER_LU, (long long)batch);
  // One factorization of the whole batch with look-ahead: the context stream runs the panel chain, the first
  // auxiliary stream the large trailing updates (fork and join with the context stream only, as everywhere
  // else, so the sequence is capturable). Fan-out 1 runs everything on the context stream.
  // the look-ahead update keeps one CU per matrix free for the panel chain that runs beside it
  const int la_wgs = c->tune.lookahead_wgs > 0 ? c->tune.lookahead_wgs
                     : c->tune.lookahead_wgs < 0 ? 0 : std::max(c->num_cus / 2, c->num_cus - batch);
  const LuLookahead la = {c->fanout >= 2 ? c->aux[0] : nullptr, c->lu_ev[0], c->lu_ev[1], la_wgs,
                          c->tune.split_tile};
  hipError_t e;
  if (precision == BEM_PREC_MIXED) {
    MatBatch<float> mb = {static_cast<cplx<float>*>(A), lda, strideA, n};
    e = lu_factor_batch<float>(mb, batch, ipiv, perm, strideP, info, nullptr, static_cast<cplx<float>*>(work),
                               c->stream, &la, c->lu_products);
  } else {
    MatBatch<double> mb = {static_cast<cplx<double>*>(A), lda, strideA, n};
    e = lu_factor_batch<double>(mb, batch, ipiv, perm, strideP, info, nullptr, static_cast<cplx<double>*>(work),
                                c->stream, &la, BEM_LU_FP32);
  }
  if (e != hipSuccess) return hipfail(e, "lu_factor launch");
  return BEM_OK;
}

int bem_lu_solve(bem_ctx* c, int precision, int n, int batch, const void* LU, long long lda, long long strideA,
                 const int* perm, long long strideP, const void* work, const void* b, long long ldb,
                 long long strideB, void* y, long long ldy, long long strideY, int nrhs, const int* skip) {
  if (!c || !LU || !perm || !b || !y || !work) return fail(BEM_E_ARG, "NULL argument");
  if (nrhs < 1 || nrhs > BEM_MAX_CHANNELS || n <= 0 || batch <= 0) return fail(BEM_E_ARG, "invalid sizes");
  ScopedTimer tm(c, BEM_TIMER_SOLVE, (long long)batch);
  hipError_t e;
  if (precision == BEM_PREC_MIXED) {
    MatBatch<float> mb = {(cplx<float>*)LU, lda, strideA, n};
    e = lu_solve_batch<float>(mb, batch, perm, strideP, (const cplx<double>*)b, ldb, strideB, (cplx<float>*)y, ldy,
                              strideY, nrhs, skip, (const cplx<float>*)work, c->stream, c->tune.trsv_wg_min_batch);
  } else if (precision == BEM_PREC_FP64) {
    MatBatch<double> mb = {(cplx<double>*)LU, lda, strideA, n};
    e = lu_solve_batch<double>(mb, batch, perm, strideP, (const cplx<double>*)b, ldb, strideB, (cplx<double>*)y,
                               ldy, strideY, nrhs, skip, (const cplx<double>*)work, c->stream,
                               c->tune.trsv_wg_min_batch);
  } else {
    return fail(BEM_E_ARG, "unknown precision");
  }
  if (e != hipSuccess) return hipfail(e, "lu_solve launch");
  return BEM_OK;
}

}  // extern "C"

static int refine_impl(bem_ctx* c, int n, int batch, const void* LU, long long lda, long long strideA,
                       const int* perm, long long strideP, const void* lu_work, const void* A0, const void* b,
                       long long ldb, long long strideB, void* x, void* work_r, void* work_y, int nrhs, int max_iter,
                       double tol, double* rel, int* iters, int* done) {
  const long long sw = lu_workspace_elems(n);
  ScopedTimer tm(c, BEM_TIMER_SOLVE, (long long)batch);
  // matrices are independent: sub-batches refine on separate streams (each solve is a chain of short launches)
  hipError_t e = fan_out(c, batch, 4, [&](int b0, int nb, hipStream_t st) -> hipError_t {
    MatBatch<float> lu = {(cplx<float>*)LU + (size_t)b0 * strideA, lda, strideA, n};
    MatBatch<float> a0 = {(cplx<float>*)A0 + (size_t)b0 * strideA, lda, strideA, n};
    const int* pm = perm + (size_t)b0 * strideP;
    const cplx<double>* bb = (const cplx<double>*)b + (size_t)b0 * strideB;
    cplx<double>* xx = (cplx<double>*)x + (size_t)b0 * strideB;
    cplx<double>* rr = (cplx<double>*)work_r + (size_t)b0 * strideB;
    cplx<float>* yy = (cplx<float>*)work_y + (size_t)b0 * strideB;
    int* dn = done + b0;
    const cplx<float>* ww = (const cplx<float>*)lu_work + (size_t)b0 * sw;
    hipError_t e2;
    // done[] = 0 with a fill kernel (no memset nodes: this sequence may be captured into a graph)
    if ((e2 = fill_int(dn, nb, 0, st))) return e2;
    if ((e2 = lu_solve_batch<float>(lu, nb, pm, strideP, bb, ldb, strideB, yy, ldb, strideB, nrhs, nullptr, ww, st,
                                    c->tune.trsv_wg_min_batch)))
      return e2;
    if ((e2 = update_batch<float>(xx, yy, ldb, ldb, strideB, strideB, n, nrhs, 1, nullptr, nb, st))) return e2;
    for (int it = 0; it <= max_iter; ++it) {
      if ((e2 = residual_batch<float>(a0, nb, bb, xx, rr, ldb, strideB, nrhs, dn, st))) return e2;
      if ((e2 = resnorm_batch(rr, bb, ldb, strideB, n, nrhs, tol, rel + b0, dn, it, iters + b0, nb, st))) return e2;
      if (it == max_iter) break;
      if ((e2 = lu_solve_batch<float>(lu, nb, pm, strideP, rr, ldb, strideB, yy, ldb, strideB, nrhs, dn, ww, st,
                                      c->tune.trsv_wg_min_batch)))
        return e2;
      if ((e2 = update_batch<float>(xx, yy, ldb, ldb, strideB, strideB, n, nrhs, 0, dn, nb, st))) return e2;
    }
    return hipSuccess;
  });
  if (e != hipSuccess) return hipfail(e, "refine");
  return BEM_OK;
}

static int refine_args_ok(bem_ctx* c, int n, int batch, const void* LU, long long lda, long long strideA,
                          const int* perm, long long strideP, const void* lu_work, const void* A0, const void* b,
                          long long ldb, long long strideB, const void* x, int nrhs, int max_iter, const double* rel,
                          const int* iters) {
  if (!c || !LU || !perm || !lu_work || !A0 || !b || !x || !rel || !iters) return fail(BEM_E_ARG, "NULL argument");
  if (n <= 0 || batch <= 0 || lda < n || strideA < lda * n || strideP < n || ldb < n || strideB < ldb * nrhs)
    return fail(BEM_E_ARG, "invalid sizes");
  if (nrhs < 1 || nrhs > BEM_MAX_CHANNELS || max_iter < 0) return fail(BEM_E_ARG, "invalid arguments");
  return BEM_OK;
}

extern "C" {

int bem_refine(bem_ctx* c, int n, int batch, const void* LU, long long lda, long long strideA, const int* perm,
               long long strideP, const void* lu_work, const void* A0, const void* b, long long ldb,
               long long strideB, void* x, void* work_r, void* work_y, int nrhs, int max_iter, double tol,
               double* rel, int* iters, int* done) {
  int rc = refine_args_ok(c, n, batch, LU, lda, strideA, perm, strideP, lu_work, A0, b, ldb, strideB, x, nrhs,
                          max_iter, rel, iters);
  if (rc) return rc;
  if (!work_r || !work_y || !done) return fail(BEM_E_ARG, "NULL argument");
  return refine_impl(c, n, batch, LU, lda, strideA, perm, strideP, lu_work, A0, b, ldb, strideB, x, work_r, work_y,
                     nrhs, max_iter, tol, rel, iters, done);
}

int bem_refine_noalloc(bem_ctx* c, int n, int batch, const void* LU, long long lda, long long strideA,
                       const int* perm, long long strideP, const void* lu_work, const void* A0, const void* b,
                       long long ldb, long long strideB, void* x, int nrhs, int max_iter, double tol, double* rel,
                       int* iters) {
  int rc = refine_args_ok(c, n, batch, LU, lda, strideA, perm, strideP, lu_work, A0, b, ldb, strideB, x, nrhs,
                          max_iter, rel, iters);
  if (rc) return rc;
  // the residuals (c128), corrections (c64) and done flags in the device's static scratch, laid out with b's
  // strides: [batch][strideB] each, 256-byte aligned
  const size_t r_bytes = ((size_t)batch * strideB * 16 + 255) / 256 * 256;
  const size_t y_bytes = ((size_t)batch * strideB * 8 + 255) / 256 * 256;
  const size_t need = r_bytes + y_bytes + (size_t)batch * sizeof(int);
  if (need > (size_t)BEM_NOALLOC_SCRATCH_BYTES)
    return fail(BEM_E_ARG, "bem_refine_noalloc needs " + std::to_string(need) + " bytes of scratch; the static "
                           "scratch holds " + std::to_string((long long)BEM_NOALLOC_SCRATCH_BYTES));
  char* base = static_cast<char*>(scratch_ptr(c->device));
  if (!base) return fail(BEM_E_STATE, "the static scratch of this device was not resolved");
  return refine_impl(c, n, batch, LU, lda, strideA, perm, strideP, lu_work, A0, b, ldb, strideB, x, base,
                     base + r_bytes, nrhs, max_iter, tol, rel, iters, reinterpret_cast<int*>(base + r_bytes + y_bytes));
}

int bem_residual(bem_ctx* c, int precision, int n, int batch, const void* A, long long lda, long long strideA,
                 const void* b, const void* x, long long ldb, long long strideB, void* work_r, int nrhs,
                 double* rel, int* scratch) {
  if (!c || !A || !b || !x || !work_r || !rel || !scratch) return fail(BEM_E_ARG, "NULL argument");
  if (nrhs < 1 || nrhs > BEM_MAX_CHANNELS || n <= 0 || batch <= 0) return fail(BEM_E_ARG, "invalid sizes");
  hipStream_t st = c->stream;
  hipError_t e;
  if (precision == BEM_PREC_MIXED) {
    MatBatch<float> mb = {(cplx<float>*)A, lda, strideA, n};
    e = residual_batch<float>(mb, batch, (const cplx<double>*)b, (const cplx<double>*)x, (cplx<double>*)work_r, ldb,
                              strideB, nrhs, nullptr, st);
  } else {
    MatBatch<double> mb = {(cplx<double>*)A, lda, strideA, n};
    e = residual_batch<double>(mb, batch, (const cplx<double>*)b, (const cplx<double>*)x, (cplx<double>*)work_r,
                               ldb, strideB, nrhs, nullptr, st);
  }
  if (e != hipSuccess) return hipfail(e, "residual");
  if ((e = fill_int(scratch, batch, 0, st)) != hipSuccess) return hipfail(e, "fill");
  e = resnorm_batch((const cplx<double>*)work_r, (const cplx<double>*)b, ldb, strideB, n, nrhs, 0.0, rel, scratch, 0,
                    scratch + batch, batch, st);
  if (e != hipSuccess) return hipfail(e, "resnorm");
  return BEM_OK;
}

int bem_growth(bem_ctx* c, int precision, int n, int batch, const void* LU, long long lda, long long strideA,
               const double* amax, double* umax, double* growth) {
  if (!c || !LU || !amax || !umax || !growth) return fail(BEM_E_ARG, "NULL argument");
  hipError_t e;
  if (precision == BEM_PREC_MIXED) {
    MatBatch<float> mb = {(cplx<float>*)LU, lda, strideA, n};
    e = growth_batch<float>(mb, batch, umax, amax, growth, c->stream);
  } else {
    MatBatch<double> mb = {(cplx<double>*)LU, lda, strideA, n};
    e = growth_batch<double>(mb, batch, umax, amax, growth, c->stream);
  }
  if (e != hipSuccess) return hipfail(e, "growth");
  return BEM_OK;
}

int bem_field(bem_ctx* c, bem_mesh* m, int precision, int nf, const double* k, int n_obs, const double* obs,
              const void* p, long long ldp, long long strideP, const void* q, long long strideQ, int n_ch,
              void* out, long long ldo, long long strideO) {
  if (!c || !m || !k || !obs || !p || !q || !out) return fail(BEM_E_ARG, "NULL argument");
  if (nf <= 0 || n_obs <= 0 || n_ch < 1 || n_ch > BEM_MAX_CHANNELS) return fail(BEM_E_ARG, "invalid sizes");
  const int ns = field_slices(m, n_obs, nf);
  const size_t cb = precision == BEM_PREC_MIXED ? 8 : 16;
  int rc;
  if ((rc = dev_alloc(m, m->wsrc, (size_t)nf * n_ch * m->nq * m->F * 2 * cb))) return rc;
  if ((rc = dev_alloc(m, m->partial, (size_t)nf * ns * n_ch * n_obs * 16))) return rc;
  note_capture_use(c, m);
  hipError_t e;
  ScopedTimer tm(c, BEM_TIMER_FIELD, (long long)nf);
  if (precision == BEM_PREC_MIXED) {
    e = field_batch<float>(m->src32, m->fv, m->srcw32, obs, m->M, m->nq * m->F, n_obs, k, (const cplx<double>*)p,
                           ldp, strideP, (const cplx<double>*)q, m->F, strideQ, n_ch, nf, (cplx<float>*)m->wsrc.p,
                           (cplx<double>*)m->partial.p, ns, (cplx<double>*)out, ldo, strideO, m->nq, c->stream);
  } else {
    e = field_batch<double>(m->src64, m->fv, m->srcw64, obs, m->M, m->nq * m->F, n_obs, k, (const cplx<double>*)p,
                            ldp, strideP, (const cplx<double>*)q, m->F, strideQ, n_ch, nf, (cplx<double>*)m->wsrc.p,
                            (cplx<double>*)m->partial.p, ns, (cplx<double>*)out, ldo, strideO, m->nq, c->stream);
  }
  if (e != hipSuccess) return hipfail(e, "field launch");
  return BEM_OK;
}

int bem_graph_begin(bem_ctx* c) {
  if (!c) return fail(BEM_E_ARG, "ctx is NULL");
  if (c->capturing) return fail(BEM_E_STATE, "already capturing");
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  c->capturing = true;
  c->capture_meshes.clear();
  return BEM_OK;
}

int bem_graph_end(bem_ctx* c, void** graph_exec) {
  if (!c || !graph_exec) return fail(BEM_E_ARG, "NULL argument");
  if (!c->capturing) return fail(BEM_E_STATE, "not capturing");
  *graph_exec = nullptr;
  hipGraph_t g = nullptr;
  c->capturing = false;
  std::vector<bem_mesh*> used;
  used.swap(c->capture_meshes);
  HIPCHK(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hipfail(e, "hipGraphInstantiate");
  bem_graph* h = nullptr;
  try {
    h = new bem_graph{ge, used};
  } catch (const std::exception&) {
    (void)hipGraphExecDestroy(ge);
    return fail(BEM_E_NOMEM, "graph handle allocation failed");
  }
  for (bem_mesh* m : h->meshes) ++m->live_graphs;
  *graph_exec = h;
  return BEM_OK;
}

int bem_graph_launch(bem_ctx* c, void* graph_exec) {
  if (!c || !graph_exec) return fail(BEM_E_ARG, "NULL argument");
  HIPCHK(hipGraphLaunch(static_cast<bem_graph*>(graph_exec)->exec, c->stream), "hipGraphLaunch");
  return BEM_OK;
}

int bem_graph_destroy(void* graph_exec) {
  if (!graph_exec) return BEM_OK;
  bem_graph* h = static_cast<bem_graph*>(graph_exec);
  (void)hipGraphExecDestroy(h->exec);
  for (bem_mesh* m : h->meshes)
    if (--m->live_graphs == 0 && m->closed) mesh_free(m);
  delete h;
  return BEM_OK;
}

int bem_mesh_live_graphs(bem_mesh* m, int* count) {
  if (!m || !count) return fail(BEM_E_ARG, "NULL argument");
  *count = m->live_graphs;
  return BEM_OK;
}

}  // extern "C"

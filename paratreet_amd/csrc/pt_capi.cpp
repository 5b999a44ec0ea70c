// C ABI of the gfx950 gravity path: context, device buffers and the stage
// sequence of one Driver::run iteration (src/Driver.h:183-288) for the
// gravity visitor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pt_gravity.h"
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kMaxBoxParts = 1024;
constexpr size_t kPinBytes = 512;  // pinned host words of the per-evaluation read-backs

enum Stage {
  S_NONE = 0,
  S_PARTICLES,
  S_UNIVERSE,
  S_KEYS,
  S_SORTED,
  S_TREE,
  S_MOMENTS,
  S_LISTS,
  S_FORCES
};

template <typename T>
struct DBuf {
  T* p = nullptr;
  int64_t cap = 0;
};

}  // namespace

struct pt_ctx {
  pt_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  int stage = S_NONE;
  bool have_forces = false;

  int64_t n = 0;
  // particle state, original order
  DBuf<double> x, y, z, vx, vy, vz, m, e, ax, ay, az, pot;
  // keys and sorted copies
  DBuf<uint64_t> keys, skeys;
  DBuf<uint32_t> perm;
  DBuf<ParticleRec> prec;  // x y z m e per particle, original order (sort gather)
  DBuf<double> xs, ys, zs, ms, es, axs, ays, azs, pots;
  DBuf<double> ustage;  // unsort staging: one 4-double record per particle
  // universe
  DBuf<double> box, bbox, box_parts;
  double h_box[6] = {0, 0, 0, 0, 0, 0};
  double soft_min = INFINITY;  // smallest particle softening (read back by pt_set_particles)
  // tree
  int64_t nn = 0, nb = 0;
  int32_t max_depth = 0;
  DBuf<uint64_t> nkey;
  DBuf<int32_t> nbegin, ncount, ndepth, ntype, nparent, nfirst;
  DBuf<int32_t> level_nodes, level_cnt;
  DBuf<int64_t> tstat;  // build read-back: {nn, nb, overflow, level counts}
  std::vector<int64_t> level_off;  // internal nodes of depth d: [level_off[d], level_off[d+1])
  int32_t split_levels = PT_MAX_DEPTH;  // split launches of the next build (last depth + 1)
  // key bits the last sort ordered (64: all; fewer: the top bits only, the
  // build then puts each leaf in full order) and those the next force
  // evaluation's sort will order (what the last tree's splits used)
  int32_t sort_bits = 64;
  int32_t sort_bits_next = 64;
  DBuf<uint32_t> flag, fpos, head, head_pos, u32_total;
  DBuf<int32_t> bucket_node;
  DBuf<double> nd;
  DBuf<int32_t> ord_h;
  DBuf<uint64_t> ord_b;
  DBuf<TreeNode> tnode;
  DBuf<GravNode> gn;
  DBuf<GravNode32> gn32;
  DBuf<Bucket> bk;
  // HEXADECAPOLE build: raw third/fourth moments (SoA), node() and open() records
  DBuf<double> hraw;
  DBuf<HexNode> hn;
  DBuf<HexOpen> ho;
  double gm = 0.0;  // GravityVisitor::monopole_gravity_factor
  DBuf<double> ewt;  // Ewald Fourier table (periodic build)
  DBuf<int32_t> ewcount;
  DBuf<unsigned long long> cpart;  // per-block counter partials
  DBuf<unsigned long long> wqueue;  // per-XCD work counters of the fused kernel
  // bucket groups of the grouped fused kernel (pt_set_fused_grouping): 0 one
  // walk per bucket, 1 automatic, 2..64 the groups' particle cap
  int32_t fused_grouping = 1;
  int32_t n_cus = 0;  // compute units of the device (0: not read yet)
  DBuf<int32_t> fg_tfirst, fg_tnb, fg_first, fg_nb;
  DBuf<uint32_t> fg_wcount, fg_woff, fg_ng;
  bool lists_valid = false;
  // lists
  DBuf<int32_t> pp_cnt, pc_cnt, pp, pc;
  DBuf<int64_t> pp_off, pc_off, i64_total;
  int64_t n_pp = 0, n_pc = 0;
  double theta = 0.7;
  double offset[3] = {0, 0, 0};
  DBuf<unsigned long long> counters;
  DBuf<int32_t> derr;
  pt_counters h_counters{};
  // leapfrog trajectory and adjoint
  DBuf<double> ckpt;      // checkpoints, 6n each (x y z vx vy vz), + final state
  DBuf<double> seg;       // positions of the steps of one segment, 3n each
  DBuf<double> adj;       // adjoint work arrays (17n)
  DBuf<double> lam_node;  // node cotangents, field-major
  // bucket groups and their member-masked lists (the adjoint's lists)
  DBuf<int32_t> gfirst, gnb, gpp_cnt, gpc_cnt, gpp, gpc;
  DBuf<uint8_t> gppm, gpcm;
  DBuf<int64_t> gpp_off, gpc_off;
  int64_t n_gpp = 0, n_gpc = 0;
  bool traj_valid = false;
  int32_t traj_steps = 0, traj_every = 1;
  int traj_mode = 0, traj_prec = 0;
  int adj_prec = PT_FP64;  // arithmetic of the tree adjoint (pt_set_adjoint_precision)
  double traj_dt = 0.0, traj_theta = 0.7;
  int64_t traj_n = 0;
  // dual-tree traversal: two frontiers of node pairs, leaf() / node() records,
  // their sorted copy, node -> bucket map, start ids, per-step state
  DBuf<uint64_t> dual_a, dual_b, dual_pp, dual_pc, dual_sorted;
  DBuf<uint32_t> dual_perm;
  DBuf<int32_t> dual_bucket_of, dual_ids;
  DBuf<unsigned long long> dual_state;
  // scratch
  DBuf<char> scratch;
  // the per-evaluation read-backs (tree sizes, counters): pinned host words
  // and an event the host polls
  void* h_pin = nullptr;
  hipEvent_t sync_ev = nullptr;
  // timing
  bool timing = false;
  hipEvent_t ev[PT_T_COUNT][2] = {};
  bool ev_used[PT_T_COUNT] = {};
  float ms_t[PT_T_COUNT] = {};
};

namespace {

pt_status fail(pt_ctx* c, pt_status st, const std::string& msg) {
  if (c) c->err = msg;
  return st;
}

pt_status hip_fail(pt_ctx* c, const char* what, hipError_t e) {
  return fail(c, PT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define PT_TRY_HIP(c, expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return hip_fail((c), #expr, _e);    \
  } while (0)

#define PT_TRY_LAUNCH(c, expr)                                              \
  do {                                                                      \
    if ((expr) != 0) {                                                      \
      hipError_t _e = hipGetLastError();                                    \
      return hip_fail((c), #expr, _e == hipSuccess ? hipErrorUnknown : _e); \
    }                                                                       \
  } while (0)

template <typename T>
hipError_t ensure(DBuf<T>& b, int64_t count) {
  if (count <= b.cap) return hipSuccess;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return e;
    b.p = nullptr;
    b.cap = 0;
  }
  int64_t c = std::max<int64_t>(count, 64);
  hipError_t e = hipMalloc((void**)&b.p, (size_t)c * sizeof(T));
  if (e != hipSuccess) return e;
  b.cap = c;
  return hipSuccess;
}

// Grows while keeping the first `keep` elements.
template <typename T>
hipError_t grow_keep(DBuf<T>& b, int64_t count, int64_t keep, hipStream_t s) {
  if (count <= b.cap) return hipSuccess;
  int64_t c = std::max<int64_t>(count, 2 * b.cap);
  T* np = nullptr;
  hipError_t e = hipMalloc((void**)&np, (size_t)c * sizeof(T));
  if (e != hipSuccess) return e;
  if (b.p && keep > 0) {
    e = hipMemcpyAsync(np, b.p, (size_t)keep * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  if (b.p) (void)hipFree(b.p);
  b.p = np;
  b.cap = c;
  return hipSuccess;
}

template <typename T>
void release(DBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

PtTree tree_of(pt_ctx* c) {
  PtTree t;
  t.key = c->nkey.p;
  t.begin = c->nbegin.p;
  t.count = c->ncount.p;
  t.depth = c->ndepth.p;
  t.type = c->ntype.p;
  t.parent = c->nparent.p;
  t.first_child = c->nfirst.p;
  return t;
}

struct Timer {
  pt_ctx* c;
  int id;
  Timer(pt_ctx* c_, int id_) : c(c_), id(id_) {
    if (c->timing) {
      (void)hipEventRecord(c->ev[id][0], c->stream);
    }
  }
  ~Timer() {
    if (c->timing) {
      (void)hipEventRecord(c->ev[id][1], c->stream);
      c->ev_used[id] = true;
    }
  }
};

// Copies `bytes` (at most kPinBytes) from the device after everything queued
// on the stream, through the pinned words, the host polling an event (a
// pageable destination would make the runtime stage the copy and block).
hipError_t read_back(pt_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c->h_pin || !c->sync_ev || bytes > kPinBytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
  }
  hipError_t e = hipMemcpyAsync(c->h_pin, src, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipEventRecord(c->sync_ev, c->stream);
  if (e != hipSuccess) return e;
  while ((e = hipEventQuery(c->sync_ev)) == hipErrorNotReady) {
  }
  if (e != hipSuccess) return e;
  memcpy(dst, c->h_pin, bytes);
  return hipSuccess;
}

pt_status check_ctx(pt_ctx* c) {
  if (!c) return PT_ERR_ARG;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, "hipSetDevice", e);
  return PT_OK;
}

pt_status need(pt_ctx* c, int stage, const char* what) {
  if (c->stage < stage) return fail(c, PT_ERR_STATE, std::string(what) + " called before its inputs exist");
  return PT_OK;
}

pt_status ensure_scratch(pt_ctx* c, int64_t bytes) {
  PT_TRY_HIP(c, ensure(c->scratch, bytes));
  return PT_OK;
}

// Node table capacity with contents preserved up to c->nn.
pt_status ensure_nodes(pt_ctx* c, int64_t count) {
  const int64_t keep = c->nn;
  PT_TRY_HIP(c, grow_keep(c->nkey, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->nbegin, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->ncount, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->ndepth, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->ntype, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->nparent, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->nfirst, count, keep, c->stream));
  PT_TRY_HIP(c, grow_keep(c->level_nodes, count, keep, c->stream));
  return PT_OK;
}

}  // namespace

extern "C" {

void pt_config_default(pt_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->min_n_subtrees = 1;
  cfg->min_n_partitions = 1;
  cfg->max_particles_per_leaf = 12;
  cfg->decomp_type = PT_DECOMP_OCT;
  cfg->tree_type = PT_TREE_OCT;
  cfg->num_iterations = 1;
  cfg->dsoft = 0.0;
  cfg->has_dsoft = 0;
  cfg->periodic = 0;
  cfg->hexadecapole = 0;
  cfg->period[0] = cfg->period[1] = cfg->period[2] = 0.0;
  cfg->n_replicas = 0;
  cfg->dual_tree = 0;
}

pt_status pt_ctx_create(const pt_config* cfg, int device, void* stream, pt_ctx** out) {
  if (!cfg || !out) return PT_ERR_ARG;
  *out = nullptr;
  if (cfg->tree_type != PT_TREE_OCT || cfg->decomp_type != PT_DECOMP_OCT) return PT_ERR_UNSUPPORTED;
  if (cfg->max_particles_per_leaf < 1) return PT_ERR_ARG;
  if (cfg->periodic) {
    // replicas of the period box (examples/Gravity.C:35-46)
    if (cfg->n_replicas < 0 || cfg->n_replicas > 16) return PT_ERR_ARG;
    for (int k = 0; k < 3; k++)
      if (!(cfg->period[k] > 0.0) || !std::isfinite(cfg->period[k])) return PT_ERR_ARG;
  }
  if (cfg->dual_tree && (cfg->periodic || cfg->hexadecapole)) return PT_ERR_UNSUPPORTED;
  pt_ctx* c = new pt_ctx();
  c->cfg = *cfg;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return PT_ERR_HIP;
  }
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
      delete c;
      return PT_ERR_HIP;
    }
    c->own_stream = true;
  }
  for (int i = 0; i < PT_T_COUNT; i++) {
    (void)hipEventCreate(&c->ev[i][0]);
    (void)hipEventCreate(&c->ev[i][1]);
  }
  if (hipHostMalloc(&c->h_pin, kPinBytes, hipHostMallocDefault) != hipSuccess) c->h_pin = nullptr;
  if (hipEventCreateWithFlags(&c->sync_ev, hipEventDisableTiming) != hipSuccess)
    c->sync_ev = nullptr;
  if (ensure(c->box, 6) != hipSuccess || ensure(c->bbox, 6) != hipSuccess ||
      ensure(c->box_parts, 6 * kMaxBoxParts) != hipSuccess || ensure(c->counters, 5) != hipSuccess ||
      ensure(c->derr, 1) != hipSuccess || ensure(c->u32_total, 4) != hipSuccess ||
      ensure(c->i64_total, 4) != hipSuccess) {
    pt_ctx_destroy(c);
    return PT_ERR_HIP;
  }
  *out = c;
  return PT_OK;
}

void pt_ctx_destroy(pt_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release(c->x); release(c->y); release(c->z); release(c->vx); release(c->vy); release(c->vz);
  release(c->m); release(c->e); release(c->ax); release(c->ay); release(c->az); release(c->pot);
  release(c->keys); release(c->skeys); release(c->perm);
  release(c->xs); release(c->ys); release(c->zs); release(c->ms); release(c->es);
  release(c->axs); release(c->ays); release(c->azs); release(c->pots); release(c->ustage);
  release(c->box); release(c->bbox); release(c->box_parts);
  release(c->nkey); release(c->nbegin); release(c->ncount); release(c->ndepth); release(c->ntype);
  release(c->nparent); release(c->nfirst); release(c->level_nodes);
  release(c->level_cnt);
  release(c->flag); release(c->fpos); release(c->head); release(c->head_pos); release(c->u32_total);
  release(c->tstat);
  release(c->prec);
  release(c->ewt);
  release(c->ewcount);
  release(c->hraw); release(c->hn); release(c->ho);
  release(c->bucket_node); release(c->nd); release(c->ord_h); release(c->ord_b);
  release(c->tnode); release(c->gn); release(c->gn32); release(c->bk); release(c->cpart);
  release(c->wqueue);
  release(c->fg_tfirst); release(c->fg_tnb); release(c->fg_first); release(c->fg_nb);
  release(c->fg_wcount); release(c->fg_woff); release(c->fg_ng);
  release(c->pp_cnt); release(c->pc_cnt); release(c->pp); release(c->pc);
  release(c->pp_off); release(c->pc_off); release(c->i64_total);
  release(c->counters); release(c->derr); release(c->scratch);
  release(c->ckpt); release(c->seg); release(c->adj); release(c->lam_node);
  release(c->gfirst); release(c->gnb); release(c->gpp_cnt); release(c->gpc_cnt);
  release(c->gpp); release(c->gpc); release(c->gppm); release(c->gpcm);
  release(c->gpp_off); release(c->gpc_off);
  release(c->dual_a); release(c->dual_b); release(c->dual_pp); release(c->dual_pc);
  release(c->dual_sorted); release(c->dual_perm); release(c->dual_bucket_of);
  release(c->dual_ids); release(c->dual_state);
  for (int i = 0; i < PT_T_COUNT; i++) {
    if (c->ev[i][0]) (void)hipEventDestroy(c->ev[i][0]);
    if (c->ev[i][1]) (void)hipEventDestroy(c->ev[i][1]);
  }
  if (c->sync_ev) (void)hipEventDestroy(c->sync_ev);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pt_last_error(const pt_ctx* c) {
  if (c) return c->err.c_str();
  const char* m = pt_noalloc_last_error();
  return m ? m : "null context";
}

pt_status pt_get_device_view(pt_ctx* c, pt_device_view* out) {
  if (!c || !out) return PT_ERR_ARG;
  out->n = c->n;
  out->x = c->x.p; out->y = c->y.p; out->z = c->z.p;
  out->vx = c->vx.p; out->vy = c->vy.p; out->vz = c->vz.p;
  out->mass = c->m.p; out->soft = c->e.p;
  out->ax = c->ax.p; out->ay = c->ay.p; out->az = c->az.p; out->pot = c->pot.p;
  out->keys = c->keys.p; out->sorted_keys = c->skeys.p; out->perm = c->perm.p;
  out->stream = (void*)c->stream;
  return PT_OK;
}

pt_status pt_set_particles(pt_ctx* c, int64_t n, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy,
                           const double* vz, const double* mass, const double* soft,
                           pt_memory where) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (n < 0 || n > 0x7fffffffll) return fail(c, PT_ERR_ARG, "particle count out of range");
  if (n > 0 && (!x || !y || !z || !mass || (!soft && !c->cfg.has_dsoft)))
    return fail(c, PT_ERR_ARG, "missing particle array");
  DBuf<double>* arrs[] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->m, &c->e,
                          &c->ax, &c->ay, &c->az, &c->pot, &c->xs, &c->ys, &c->zs,
                          &c->ms, &c->es, &c->axs, &c->ays, &c->azs, &c->pots};
  for (auto* a : arrs) PT_TRY_HIP(c, ensure(*a, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->keys, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->skeys, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->perm, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->head, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->head_pos, std::max<int64_t>(n, 1)));
  PT_TRY_HIP(c, ensure(c->bucket_node, std::max<int64_t>(n, 1)));
  c->n = n;
  c->have_forces = false;
  c->traj_valid = false;
  const hipMemcpyKind kind = where == PT_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const size_t bytes = (size_t)n * sizeof(double);
  if (n > 0) {
    PT_TRY_HIP(c, hipMemcpyAsync(c->x.p, x, bytes, kind, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(c->y.p, y, bytes, kind, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(c->z.p, z, bytes, kind, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(c->m.p, mass, bytes, kind, c->stream));
    const double* vv[3] = {vx, vy, vz};
    double* vd[3] = {c->vx.p, c->vy.p, c->vz.p};
    for (int k = 0; k < 3; k++) {
      if (vv[k]) {
        PT_TRY_HIP(c, hipMemcpyAsync(vd[k], vv[k], bytes, kind, c->stream));
      } else {
        PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)vd[k], n, 0ull, c->stream));
      }
    }
    if (c->cfg.has_dsoft) {
      // Reader::setSoft: every particle gets dSoft
      unsigned long long bits;
      std::memcpy(&bits, &c->cfg.dsoft, sizeof(bits));
      PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->e.p, n, bits, c->stream));
    } else {
      PT_TRY_HIP(c, hipMemcpyAsync(c->e.p, soft, bytes, kind, c->stream));
    }
    PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->ax.p, n, 0ull, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->ay.p, n, 0ull, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->az.p, n, 0ull, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->pot.p, n, 0ull, c->stream));
    double* smin = c->box_parts.p + 6 * kMaxBoxParts - 1;
    PT_TRY_LAUNCH(c, pt_launch_min(c->e.p, n, c->box_parts.p, kMaxBoxParts, smin, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(&c->soft_min, smin, sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
  } else {
    c->soft_min = INFINITY;
  }
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  c->stage = S_PARTICLES;
  return PT_OK;
}

pt_status pt_remake_universe(pt_ctx* c, double box_out[6]) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_PARTICLES, "pt_remake_universe"))) return st;
  if (c->n == 0) return fail(c, PT_ERR_ARG, "no particles");
  {
    Timer t(c, PT_T_UNIVERSE);
    PT_TRY_LAUNCH(c, pt_launch_universe(c->x.p, c->y.p, c->z.p, c->n, c->box_parts.p, kMaxBoxParts,
                                        c->bbox.p, c->box.p, c->stream));
  }
  if (box_out) {
    PT_TRY_HIP(c, hipMemcpyAsync(c->h_box, c->box.p, 6 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(box_out, c->h_box, sizeof(c->h_box));
  }
  c->stage = S_UNIVERSE;
  return PT_OK;
}

pt_status pt_assign_keys(pt_ctx* c) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_UNIVERSE, "pt_assign_keys"))) return st;
  Timer t(c, PT_T_KEYS);
  PT_TRY_HIP(c, ensure(c->prec, std::max<int64_t>(c->n, 1)));
  PT_TRY_LAUNCH(c, pt_launch_keys(c->x.p, c->y.p, c->z.p, c->m.p, c->e.p, c->n, c->box.p,
                                  c->keys.p, c->prec.p, c->stream));
  c->stage = S_KEYS;
  return PT_OK;
}

}  // extern "C"

namespace {

// Sorts on the top `bits` key bits (64: the full key, then the sorted copies
// of the particles are gathered; fewer: pt_build_tree finishes the order per
// leaf and gathers).
pt_status sort_stage(pt_ctx* c, int bits) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_KEYS, "pt_sort"))) return st;
  if ((st = ensure_scratch(c, pt_sort_scratch_bytes(c->n)))) return st;
  Timer t(c, PT_T_SORT);
  PT_TRY_LAUNCH(c, pt_radix_sort(c->keys.p, c->skeys.p, c->perm.p, c->n, 64 - bits, 64,
                                 c->scratch.p, c->stream));
  c->sort_bits = bits;
  if (bits >= 64)
    PT_TRY_LAUNCH(c, pt_launch_gather(c->perm.p, c->n, c->prec.p, c->xs.p, c->ys.p, c->zs.p,
                                      c->ms.p, c->es.p, c->stream));
  c->stage = S_SORTED;
  return PT_OK;
}

// pt_build_tree's answer when a top-bits sort did not cover the tree's
// splits (internal; the caller sorts the full keys and builds again).
constexpr pt_status kRedoFullSort = (pt_status)1;

}  // namespace

extern "C" {

pt_status pt_sort(pt_ctx* c) { return sort_stage(c, 64); }

// Level-parallel build driven from the device: one split launch per level
// (internal-node counts stay on the device), a host check every 4 levels, and
// a rebuild with larger tables if a capacity was exceeded.
pt_status pt_build_tree(pt_ctx* c) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_SORTED, "pt_build_tree"))) return st;
  const int B = c->cfg.max_particles_per_leaf;
  Timer t(c, PT_T_BUILD);
  // internal nodes of one level are disjoint and hold more than B particles
  const int64_t per_level = c->n / (B + 1) + 1;
  int64_t level_cap = std::max<int64_t>(64, 2 * per_level);
  PT_TRY_HIP(c, ensure(c->level_cnt, PT_MAX_DEPTH + 2));
  int capacity_retries = 0;
  for (int attempt = 0;; attempt++) {
    const int64_t node_cap = 1 + 8 * level_cap;
    c->nn = 0;
    if ((st = ensure_nodes(c, node_cap))) return st;
    PT_TRY_HIP(c, ensure(c->level_nodes, level_cap));
    PtTree tr = tree_of(c);
    PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->derr.p, 1, 0u, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_root(c->n, B, tr, c->level_cnt.p, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->level_nodes.p, 1, 0u, c->stream));
    // the splits of the previous build's depth plus one level (all levels the
    // first time; a deeper tree is detected below and rebuilt with all);
    // the buckets follow on the device and one read-back returns counts,
    // sizes and the overflow flag
    const int levels = c->split_levels;
    int64_t bound = 1;  // upper bound of the level's internal-node count
    for (int depth = 0; depth < levels; depth++) {
      PT_TRY_LAUNCH(c, pt_launch_split_level(c->skeys.p, c->level_nodes.p, c->level_cnt.p, depth,
                                             bound, B, tr, node_cap, level_cap, c->derr.p,
                                             c->stream));
      bound = std::min<int64_t>(bound * 8, per_level);
    }
    if ((st = ensure_scratch(c, pt_scan_scratch_bytes(c->n)))) return st;
    PT_TRY_HIP(c, ensure(c->tstat, 3 + PT_MAX_DEPTH + 2));
    PT_TRY_LAUNCH(c, pt_launch_buckets(tr, node_cap, c->n, levels, c->level_cnt.p, c->derr.p,
                                       c->head.p, c->head_pos.p, c->scratch.p, c->u32_total.p,
                                       c->bucket_node.p, c->tstat.p, c->stream));
    if (c->sort_bits < 64) {
      // top-bits sort: each leaf's particles into full key order, then the
      // sorted copies (idempotent if a retry below builds again)
      PT_TRY_LAUNCH(c, pt_launch_leaf_fixup(tr, c->n, c->bucket_node.p, c->tstat.p, c->skeys.p,
                                            c->perm.p, c->stream));
      PT_TRY_LAUNCH(c, pt_launch_gather(c->perm.p, c->n, c->prec.p, c->xs.p, c->ys.p, c->zs.p,
                                        c->ms.p, c->es.p, c->stream));
    }
    int64_t stat[3 + PT_MAX_DEPTH + 2] = {0};
    PT_TRY_HIP(c, read_back(c, stat, c->tstat.p, sizeof(stat)));
    if (c->sort_bits < 64) {
      // valid when every split compared only sorted bits: a split creating
      // depth d + 1 reads the top 3 (d + 1) + 1 bits
      int dmax = 0;
      for (int d = 0; d <= PT_MAX_DEPTH; d++)
        if (stat[3 + d] > 0) dmax = d + 1;
      if (stat[2] || stat[0] == 0 || 3 * dmax + 1 > c->sort_bits) return kRedoFullSort;
    }
    if (stat[2]) {
      if (++capacity_retries > 3) return fail(c, PT_ERR_CAPACITY, "octree node table overflow");
      level_cap *= 4;
      continue;
    }
    if (stat[0] == 0) {
      // deeper than predicted: all levels
      if (levels >= PT_MAX_DEPTH || attempt > 8)
        return fail(c, PT_ERR_CAPACITY, "octree build did not complete");
      c->split_levels = PT_MAX_DEPTH;
      continue;
    }
    // level table: internal nodes of depth d are [level_off[d], level_off[d+1])
    const int64_t* cnt = stat + 3;
    c->level_off.assign(1, 0);
    c->max_depth = 0;
    int64_t internal = 0;
    for (int d = 0; d <= PT_MAX_DEPTH; d++) {
      internal += cnt[d];
      c->level_off.push_back(internal);
      if (cnt[d] > 0) c->max_depth = d + 1;
    }
    c->nn = stat[0];
    c->nb = stat[1];
    c->split_levels = std::min<int32_t>(PT_MAX_DEPTH, c->max_depth + 1);
    break;
  }
  c->stage = S_TREE;
  return PT_OK;
}

}  // extern "C"

namespace {

// Sort + build of a force evaluation.  The sort orders only the key bits the
// previous tree's splits used (the first evaluation, buckets above kFixupMax
// and deeper trees use all 64, the last after a failed check); the result is
// the full sort either way.
pt_status sort_and_build(pt_ctx* c) {
  pt_status st;
  const int bits = c->cfg.max_particles_per_leaf <= kFixupMax ? c->sort_bits_next : 64;
  if ((st = sort_stage(c, bits))) return st;
  st = pt_build_tree(c);
  if (st == kRedoFullSort) {
    if ((st = sort_stage(c, 64))) return st;
    st = pt_build_tree(c);
  }
  if (st) return st;
  c->sort_bits_next = std::min(64, ((3 * c->max_depth + 1 + 7) / 8) * 8);
  return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_upward(pt_ctx* c) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_TREE, "pt_upward"))) return st;
  Timer t(c, PT_T_MOMENTS);
  const int64_t nn = c->nn;
  PT_TRY_HIP(c, ensure(c->nd, PT_CD_NFIELDS * nn));
  PT_TRY_HIP(c, ensure(c->ord_h, nn));
  PT_TRY_HIP(c, ensure(c->ord_b, nn));
  PT_TRY_HIP(c, ensure(c->tnode, nn));
  PT_TRY_HIP(c, ensure(c->gn, nn));
  PT_TRY_HIP(c, ensure(c->gn32, nn));
  PT_TRY_HIP(c, ensure(c->bk, std::max<int64_t>(c->nb, 1)));
  PtTree tr = tree_of(c);
  const PtPackOut po{c->tnode.p, c->gn.p, c->gn32.p, c->bk.p, c->head_pos.p};
  PT_TRY_LAUNCH(c, pt_launch_leaf_moments(tr, nn, c->xs.p, c->ys.p, c->zs.p, c->ms.p, c->es.p,
                                          c->nd.p, po, c->stream));
  // internal nodes, deepest level first: a level of more than kRunNodes
  // nodes gets its own launch, consecutive smaller ones share one work-group
  constexpr int64_t kRunNodes = 64;
  const int nlev = (int)c->level_off.size() - 1;
  PtLevelRun run{};
  for (int d = nlev - 1; d >= 0; d--) {
    const int64_t a = c->level_off[d], b = c->level_off[d + 1];
    if (b <= a) continue;
    if (b - a <= kRunNodes && run.nlev < PT_MAX_RUN) {
      run.off[run.nlev] = (int32_t)a;
      run.cnt[run.nlev] = (int32_t)(b - a);
      run.nlev++;
      continue;
    }
    PT_TRY_LAUNCH(c, pt_launch_upward_run(tr, c->level_nodes.p, run, c->nd.p, nn, c->ord_h.p,
                                          c->ord_b.p, po, c->stream));
    run.nlev = 0;
    PT_TRY_LAUNCH(c, pt_launch_upward_level(tr, c->level_nodes.p + a, (int32_t)(b - a), c->nd.p,
                                            nn, c->ord_h.p, c->ord_b.p, po, c->stream));
  }
  PT_TRY_LAUNCH(c, pt_launch_upward_run(tr, c->level_nodes.p, run, c->nd.p, nn, c->ord_h.p,
                                        c->ord_b.p, po, c->stream));
  if (c->cfg.hexadecapole) {
    // third and fourth moments once every centre of mass is final
    PT_TRY_HIP(c, ensure(c->hraw, (int64_t)PT_HEXA_RAW * nn));
    PT_TRY_HIP(c, ensure(c->hn, nn));
    PT_TRY_HIP(c, ensure(c->ho, nn));
    PT_TRY_LAUNCH(c, pt_launch_hexa_leaf(tr, c->bucket_node.p, c->nb, c->xs.p, c->ys.p, c->zs.p,
                                         c->ms.p, c->nd.p, nn, c->hraw.p, c->stream));
    for (int d = nlev - 1; d >= 0; d--) {
      const int64_t a = c->level_off[d], b = c->level_off[d + 1];
      if (b > a)
        PT_TRY_LAUNCH(c, pt_launch_hexa_level(tr, c->level_nodes.p + a, (int32_t)(b - a),
                                              c->nd.p, nn, c->hraw.p, c->stream));
    }
    PT_TRY_LAUNCH(c, pt_launch_hexa_pack(tr, c->nd.p, c->hraw.p, nn, c->hn.p, c->ho.p,
                                         c->stream));
  }
  c->stage = S_MOMENTS;
  c->lists_valid = false;
  return PT_OK;
}

static pt_status set_visitor(pt_ctx* c, double theta, const double offset[3], double* gf) {
  if (!(theta > 0.0)) return fail(c, PT_ERR_ARG, "theta must be positive");
  c->theta = theta;
  for (int k = 0; k < 3; k++) c->offset[k] = offset ? offset[k] : 0.0;
  // GravityVisitor(offset, theta) (examples/GravityVisitor.h:16-19)
  const double opening_geometry_factor_squared = 4.0 / 3.0;
  *gf = opening_geometry_factor_squared / (theta * theta);
  c->gm = std::sqrt(opening_geometry_factor_squared) / (theta * theta * theta * theta);
  return PT_OK;
}

// Reads back the reduced counters and the overflow flag (synchronises).
// Reduces the counter partials (with the overflow flag alongside) and reads
// both back with one synchronise; with `unsort` the results are first put
// back in caller order (the fused path).
static pt_status finish_counters(pt_ctx* c, int64_t nblocks, int32_t* herr, bool unsort = false,
                                 bool accumulate = false) {
  PT_TRY_LAUNCH(c, pt_launch_reduce_counters(c->cpart.p, nblocks, c->counters.p, c->derr.p,
                                             c->stream, accumulate));
  if (unsort) {
    PT_TRY_HIP(c, ensure(c->ustage, 4 * c->n));
    PT_TRY_LAUNCH(c, pt_launch_unsort(c->perm.p, c->n, c->axs.p, c->ays.p, c->azs.p, c->pots.p,
                                      c->ax.p, c->ay.p, c->az.p, c->pot.p, c->ustage.p,
                                      c->stream));
  }
  unsigned long long h[5];
  PT_TRY_HIP(c, read_back(c, h, c->counters.p, sizeof(h)));
  c->h_counters.n_node_ints = h[0];
  c->h_counters.n_part_ints = h[1];
  c->h_counters.n_opens = h[2];
  c->h_counters.n_closes = h[3];
  *herr = (int32_t)h[4];
  return PT_OK;
}

pt_status pt_traverse(pt_ctx* c, double theta, const double offset[3], pt_counters* counters) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_MOMENTS, "pt_traverse"))) return st;
  double gf;
  if ((st = set_visitor(c, theta, offset, &gf))) return st;
  const int64_t nb = c->nb;
  {
    Timer t(c, PT_T_TRAVERSE);
    const int64_t nblk = pt_traverse_blocks(nb);
    PT_TRY_HIP(c, ensure(c->pp_cnt, std::max<int64_t>(nb, 1)));
    PT_TRY_HIP(c, ensure(c->pc_cnt, std::max<int64_t>(nb, 1)));
    PT_TRY_HIP(c, ensure(c->pp_off, nb + 1));
    PT_TRY_HIP(c, ensure(c->pc_off, nb + 1));
    PT_TRY_HIP(c, ensure(c->cpart, std::max<int64_t>(nblk, 1) * 4));
    if ((st = ensure_scratch(c, pt_scan_scratch_bytes(nb + 1)))) return st;
    PT_TRY_LAUNCH(c, pt_launch_fill_u64(c->cpart.p, std::max<int64_t>(nblk, 1) * 4, 0ull, c->stream));
    PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->derr.p, 1, 0u, c->stream));
    const HexOpen* ho = c->cfg.hexadecapole ? c->ho.p : nullptr;
    PT_TRY_LAUNCH(c, pt_launch_traverse(false, c->tnode.p, ho, c->bk.p, nb, gf, c->gm, c->offset,
                                        c->pp_cnt.p, c->pc_cnt.p, nullptr, nullptr, nullptr,
                                        nullptr, c->cpart.p, c->derr.p, c->stream));
    PT_TRY_LAUNCH(c, pt_exclusive_scan_i32_to_i64(c->pp_cnt.p, c->pp_off.p, nb, c->scratch.p,
                                                  c->pp_off.p + nb, c->stream));
    PT_TRY_LAUNCH(c, pt_exclusive_scan_i32_to_i64(c->pc_cnt.p, c->pc_off.p, nb, c->scratch.p,
                                                  c->pc_off.p + nb, c->stream));
    int64_t tot[2] = {0, 0};
    PT_TRY_HIP(c, read_back(c, &tot[0], c->pp_off.p + nb, 8));
    PT_TRY_HIP(c, read_back(c, &tot[1], c->pc_off.p + nb, 8));
    int32_t herr = 0;
    if ((st = finish_counters(c, std::max<int64_t>(nblk, 1), &herr))) return st;
    if (herr) return fail(c, PT_ERR_CAPACITY, "traversal stack overflow (PT_WALK_STACK)");
    c->n_pp = tot[0];
    c->n_pc = tot[1];
    PT_TRY_HIP(c, ensure(c->pp, std::max<int64_t>(c->n_pp, 1)));
    PT_TRY_HIP(c, ensure(c->pc, std::max<int64_t>(c->n_pc, 1)));
    PT_TRY_LAUNCH(c, pt_launch_traverse(true, c->tnode.p, ho, c->bk.p, nb, gf, c->gm, c->offset,
                                        nullptr, nullptr, c->pp_off.p, c->pc_off.p, c->pp.p,
                                        c->pc.p, c->cpart.p, c->derr.p, c->stream));
  }
  if (counters) *counters = c->h_counters;
  c->lists_valid = true;
  return PT_OK;
}

// The compensated-FP32 pair terms grow as powers of 1/softening (the
// softened SPLINEQ terms to (2/twoh)^7 in the forward, their r-derivatives to
// (2/twoh)^8 in the adjoint) and reach the FP32 range (3.4e38) when a
// softening is below about 2e-5 length units: the FP32C forms refuse such
// inputs instead of returning infinities.  FP64 has no such limit.
constexpr double kFp32cMinSoft = 2.5e-5;

static pt_status check_fp32c_range(pt_ctx* c) {
  if (!(c->soft_min >= kFp32cMinSoft)) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "PT_FP32C needs every softening >= %g (the smallest is %g): its FP32 pair terms "
                  "grow as softening^-8; use PT_FP64, or lengths in larger units",
                  kFp32cMinSoft, c->soft_min);
    return fail(c, PT_ERR_UNSUPPORTED, msg);
  }
  return PT_OK;
}

// node() / leaf() over the context's lists into the sorted result arrays
// (written, not added).
static pt_status launch_list_interact(pt_ctx* c, pt_precision precision) {
  if (precision != PT_FP64) {
    pt_status st = check_fp32c_range(c);
    if (st) return st;
  }
  Timer t(c, PT_T_FORCE);
  if (c->cfg.hexadecapole) {
    PT_TRY_LAUNCH(c, pt_launch_interact_hexa(c->bk.p, c->nb, c->tnode.p, c->hn.p, c->xs.p, c->ys.p,
                                             c->zs.p, c->ms.p, c->es.p, c->pp_off.p, c->pp.p,
                                             c->pc_off.p, c->pc.p, c->offset, c->axs.p, c->ays.p,
                                             c->azs.p, c->pots.p, c->stream));
  } else {
    PT_TRY_LAUNCH(c, pt_launch_interact((int)precision, c->bk.p, c->nb, c->tnode.p, c->gn.p,
                                        c->gn32.p, c->xs.p, c->ys.p, c->zs.p, c->ms.p, c->es.p,
                                        c->pp_off.p, c->pp.p, c->pc_off.p, c->pc.p, c->offset,
                                        c->axs.p, c->ays.p, c->azs.p, c->pots.p, c->stream));
  }
  return PT_OK;
}

pt_status pt_interact(pt_ctx* c, pt_precision precision) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->lists_valid) return fail(c, PT_ERR_STATE, "pt_interact called before pt_traverse");
  if (precision != PT_FP64 && precision != PT_FP32C) return fail(c, PT_ERR_ARG, "bad precision");
  if (c->cfg.hexadecapole && precision != PT_FP64)
    return fail(c, PT_ERR_UNSUPPORTED,
                "the hexadecapole list path evaluates node() in FP64 only (pt_compute_forces "
                "has the compensated-FP32 form)");
  if ((st = launch_list_interact(c, precision))) return st;
  PT_TRY_HIP(c, ensure(c->ustage, 4 * c->n));
  PT_TRY_LAUNCH(c, pt_launch_unsort(c->perm.p, c->n, c->axs.p, c->ays.p, c->azs.p, c->pots.p,
                                    c->ax.p, c->ay.p, c->az.p, c->pot.p, c->ustage.p, c->stream));
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  c->stage = S_FORCES;
  c->have_forces = true;
  return PT_OK;
}

// Frontier items and records of one dual-tree traversal at most (beyond it
// the pair walk, which never prunes below two internal nodes whose boxes do
// not enclose one another, is reported as a capacity error).
constexpr int64_t kDualMaxItems = int64_t(1) << 28;

// DualTraverser::doTrav from the pair of nodes with keys (skey, tkey), one
// launch per depth of the pair walk: leaves the leaf() / node() calls as the
// context's per-bucket lists and the counters in c->h_counters.
static pt_status dual_lists(pt_ctx* c, double theta, const double offset[3], uint64_t skey,
                            uint64_t tkey) {
  pt_status st;
  if ((st = need(c, S_MOMENTS, "pt_start_dual"))) return st;
  if (c->cfg.periodic || c->cfg.hexadecapole)
    return fail(c, PT_ERR_UNSUPPORTED,
                "the dual-tree traversal runs on non-periodic quadrupole contexts "
                "(examples/Gravity.C:30 aborts on dual_tree with periodic)");
  double gf;
  if ((st = set_visitor(c, theta, offset, &gf))) return st;
  const int64_t nn = c->nn, nb = c->nb;
  Timer tm(c, PT_T_TRAVERSE);
  PT_TRY_HIP(c, ensure(c->dual_ids, 2));
  PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->dual_ids.p, 2, 0xffffffffu, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_dual_find(c->nkey.p, nn, skey, tkey, c->dual_ids.p, c->stream));
  int32_t ids[2] = {-1, -1};
  PT_TRY_HIP(c, hipMemcpyAsync(ids, c->dual_ids.p, sizeof(ids), hipMemcpyDeviceToHost, c->stream));
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  if (ids[0] < 0 || ids[1] < 0) return fail(c, PT_ERR_ARG, "pt_start_dual: no node with that key");
  PT_TRY_HIP(c, ensure(c->dual_bucket_of, std::max<int64_t>(nn, 1)));
  PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->dual_bucket_of.p, nn, 0xffffffffu, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_dual_bucket_of(c->bk.p, nb, c->dual_bucket_of.p, c->stream));
  PT_TRY_HIP(c, ensure(c->dual_a, std::max<int64_t>(4 * nn, 1024)));
  PT_TRY_HIP(c, ensure(c->dual_b, std::max<int64_t>(4 * nn, 1024)));
  PT_TRY_HIP(c, ensure(c->dual_pp, std::max<int64_t>(8 * nb, 1024)));
  PT_TRY_HIP(c, ensure(c->dual_pc, std::max<int64_t>(8 * nb, 1024)));
  PT_TRY_HIP(c, ensure(c->dual_state, 8));
  const unsigned long long seed =
      ((unsigned long long)(uint32_t)ids[0] << 32) | (unsigned long long)(uint32_t)ids[1];
  PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)c->dual_a.p, 1, seed, c->stream));
  int64_t n_cur = 1, n_pp = 0, n_pc = 0;
  unsigned long long tot[4] = {0, 0, 0, 0};
  for (int level = 0; n_cur > 0; level++) {
    if (level > 4 * PT_MAX_DEPTH + 8)
      return fail(c, PT_ERR_CAPACITY, "dual-tree traversal did not finish");
    unsigned long long h[7];
    for (int retry = 0;; retry++) {
      unsigned long long init[7] = {0, (unsigned long long)n_pp, (unsigned long long)n_pc, 0, 0,
                                    0, 0};
      PT_TRY_HIP(c, hipMemcpyAsync(c->dual_state.p, init, sizeof(init), hipMemcpyHostToDevice,
                                   c->stream));
      PT_TRY_LAUNCH(c, pt_launch_dual_step(c->dual_a.p, n_cur, c->tnode.p, c->nd.p, c->ndepth.p,
                                           c->dual_bucket_of.p, nn, gf, c->offset, c->dual_b.p,
                                           c->dual_b.cap, c->dual_pp.p, c->dual_pp.cap,
                                           c->dual_pc.p, c->dual_pc.cap, c->dual_state.p,
                                           c->stream));
      PT_TRY_HIP(c, hipMemcpyAsync(h, c->dual_state.p, sizeof(h), hipMemcpyDeviceToHost,
                                   c->stream));
      PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
      const int64_t want_next = (int64_t)h[0], want_pp = (int64_t)h[1], want_pc = (int64_t)h[2];
      if (want_next <= c->dual_b.cap && want_pp <= c->dual_pp.cap && want_pc <= c->dual_pc.cap)
        break;
      if (retry > 0 || want_next > kDualMaxItems || want_pp > kDualMaxItems ||
          want_pc > kDualMaxItems)
        return fail(c, PT_ERR_CAPACITY, "dual-tree traversal exceeds its pair capacity");
      // grow what overflowed (records keep what earlier depths wrote) and
      // repeat this depth
      PT_TRY_HIP(c, ensure(c->dual_b, want_next));
      PT_TRY_HIP(c, grow_keep(c->dual_pp, want_pp, n_pp, c->stream));
      PT_TRY_HIP(c, grow_keep(c->dual_pc, want_pc, n_pc, c->stream));
    }
    for (int k = 0; k < 4; k++) tot[k] += h[3 + k];
    n_pp = (int64_t)h[1];
    n_pc = (int64_t)h[2];
    n_cur = (int64_t)h[0];
    std::swap(c->dual_a, c->dual_b);
  }
  // records -> per-bucket lists in (bucket, source) order
  int end_bit = 32;
  while (end_bit < 64 && (int64_t(1) << (end_bit - 32)) <= nb) end_bit++;
  const int64_t nrec = std::max<int64_t>(std::max(n_pp, n_pc), 1);
  PT_TRY_HIP(c, ensure(c->dual_sorted, nrec));
  PT_TRY_HIP(c, ensure(c->dual_perm, nrec));
  if ((st = ensure_scratch(c, pt_sort_scratch_bytes(nrec)))) return st;
  PT_TRY_HIP(c, ensure(c->pp, std::max<int64_t>(n_pp, 1)));
  PT_TRY_HIP(c, ensure(c->pc, std::max<int64_t>(n_pc, 1)));
  PT_TRY_HIP(c, ensure(c->pp_off, nb + 1));
  PT_TRY_HIP(c, ensure(c->pc_off, nb + 1));
  PT_TRY_LAUNCH(c, pt_radix_sort(c->dual_pp.p, c->dual_sorted.p, c->dual_perm.p, n_pp, 0, end_bit,
                                 c->scratch.p, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_records_to_lists(c->dual_sorted.p, n_pp, nb, c->pp.p, c->pp_off.p,
                                              c->stream));
  PT_TRY_LAUNCH(c, pt_radix_sort(c->dual_pc.p, c->dual_sorted.p, c->dual_perm.p, n_pc, 0, end_bit,
                                 c->scratch.p, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_records_to_lists(c->dual_sorted.p, n_pc, nb, c->pc.p, c->pc_off.p,
                                              c->stream));
  c->n_pp = n_pp;
  c->n_pc = n_pc;
  c->lists_valid = true;
  c->h_counters.n_node_ints = tot[0];
  c->h_counters.n_part_ints = tot[1];
  c->h_counters.n_opens = tot[2];
  c->h_counters.n_closes = tot[3];
  return PT_OK;
}

pt_status pt_start_dual(pt_ctx* c, double theta, const double offset[3], uint64_t source_key,
                        uint64_t target_key, pt_precision precision, pt_counters* counters) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (precision != PT_FP64 && precision != PT_FP32C) return fail(c, PT_ERR_ARG, "bad precision");
  if ((st = dual_lists(c, theta, offset, source_key, target_key))) return st;
  const pt_counters k = c->h_counters;
  if ((st = pt_interact(c, precision))) return st;
  c->h_counters = k;
  if (counters) *counters = k;
  return PT_OK;
}

// Fused walk + interactions (no lists), with a large-stack rerun if the
// small walk stack (192 entries) overflowed for some bucket.
// The particle cap of the grouped walk's bucket groups, 0 for one walk per
// bucket.  Automatic: the largest of 64, 32 and 16 that still gives every
// resident wave slot (four per SIMD) about six groups, 16 for smaller sets
// down to one group per slot, and one walk per bucket below that (larger
// groups share more of the walk, more groups keep the device busy to the
// end).  Measured on 100k, 256k and 1M Plummer spheres and the 16M halo
// mixture, this picks the fastest of those caps in each case.
static int group_cap(pt_ctx* c) {
  if (c->fused_grouping == 0) return 0;
  if (c->fused_grouping >= 2) return c->fused_grouping;
  if (c->n_cus == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess ||
        cus <= 0)
      cus = 256;
    c->n_cus = cus;
  }
  const int64_t slots = (int64_t)c->n_cus * 16;
  for (int cap = 64; cap >= 32; cap /= 2)
    if (c->n >= 6 * cap * slots) return cap;
  return c->n >= 16 * slots ? 16 : 0;
}

static pt_status gravity_fused(pt_ctx* c, double theta, const double offset[3],
                               pt_precision precision, pt_counters* counters) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_MOMENTS, "pt_compute_forces"))) return st;
  if (precision != PT_FP64 && precision != PT_FP32C) return fail(c, PT_ERR_ARG, "bad precision");
  double gf;
  if ((st = set_visitor(c, theta, offset, &gf))) return st;
  if (precision != PT_FP64 && (st = check_fp32c_range(c))) return st;
  const int64_t nb = c->nb;
  const bool hexa = c->cfg.hexadecapole != 0;
  const int cap = group_cap(c);
  const bool grouped = cap > 0;
  const int64_t nblk = std::max<int64_t>(
      grouped ? pt_grouped_gravity_blocks(nb, c->n, cap)
              : (hexa ? pt_hexa_gravity_blocks(nb) : pt_gravity_blocks(nb, (int)precision)),
      1);
  PT_TRY_HIP(c, ensure(c->cpart, nblk * 4));
  PT_TRY_HIP(c, ensure(c->wqueue, 8));
  if (grouped) {
    // the bucket groups of this tree (every replica walks with them)
    const int64_t nw = pt_fgroup_windows(nb);
    PT_TRY_HIP(c, ensure(c->fg_tfirst, nb));
    PT_TRY_HIP(c, ensure(c->fg_tnb, nb));
    PT_TRY_HIP(c, ensure(c->fg_first, nb));
    PT_TRY_HIP(c, ensure(c->fg_nb, nb));
    PT_TRY_HIP(c, ensure(c->fg_wcount, nw));
    PT_TRY_HIP(c, ensure(c->fg_woff, nw));
    PT_TRY_HIP(c, ensure(c->fg_ng, 1));
    if ((st = ensure_scratch(c, pt_scan_scratch_bytes(nw)))) return st;
    Timer t(c, PT_T_TRAVERSE);
    PT_TRY_LAUNCH(c, pt_launch_fgroups(c->bk.p, nb, cap, c->fg_tfirst.p, c->fg_tnb.p, c->fg_wcount.p,
                                       c->fg_woff.p, c->scratch.p, c->fg_ng.p, c->fg_first.p,
                                       c->fg_nb.p, c->stream));
  }
  // the visitor offsets: one, or the (2N+1)^3 period replicas in the order of
  // ExMain::traversalFn (examples/Gravity.C:35-46), each added to `offset`;
  // every replica's traversal adds to the accelerations and the counters
  std::vector<std::array<double, 3>> offs;
  const int nr = c->cfg.periodic ? c->cfg.n_replicas : 0;
  for (int X = -nr; X <= nr; X++)
    for (int Y = -nr; Y <= nr; Y++)
      for (int Z = -nr; Z <= nr; Z++)
        offs.push_back({c->offset[0] + X * c->cfg.period[0], c->offset[1] + Y * c->cfg.period[1],
                        c->offset[2] + Z * c->cfg.period[2]});
  // examples/Gravity.C:30-33: with dual_tree, Subtree::startDual's traversal
  // first, then the down traversal adds to the same accelerations
  const bool dual = c->cfg.dual_tree != 0;
  pt_counters dual_cnt{};
  if (dual) {
    if ((st = dual_lists(c, theta, offset, 1, 1))) return st;
    dual_cnt = c->h_counters;
    if ((st = set_visitor(c, theta, offset, &gf))) return st;
  }
  for (int attempt = 0; attempt < 2; attempt++) {
    PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->derr.p, 1, 0u, c->stream));
    if (dual && (st = launch_list_interact(c, precision))) return st;
    for (size_t r = 0; r < offs.size(); r++) {
      const bool acc = r > 0 || dual;
      const double* off = offs[r].data();
      PT_TRY_LAUNCH(c, pt_launch_fill_u64(c->cpart.p, nblk * 4, 0ull, c->stream));
      {
        Timer t(c, PT_T_FORCE);
        if (grouped)
          PT_TRY_LAUNCH(c, pt_launch_grouped_gravity(
                               (int)precision, attempt == 1, c->bk.p, c->fg_first.p, c->fg_nb.p,
                               c->fg_ng.p, nb, c->n, cap, c->tnode.p, c->gn.p, c->gn32.p,
                               hexa ? c->ho.p : nullptr, hexa ? c->hn.p : nullptr, c->xs.p,
                               c->ys.p, c->zs.p, c->ms.p, c->es.p, gf, c->gm, off, c->axs.p,
                               c->ays.p, c->azs.p, c->pots.p, c->cpart.p, c->derr.p, acc,
                               c->stream));
        else if (hexa)
          PT_TRY_LAUNCH(c, pt_launch_hexa_gravity((int)precision, attempt == 1, c->bk.p, nb,
                                                  c->tnode.p, c->ho.p,
                                                  c->hn.p, c->xs.p, c->ys.p, c->zs.p, c->ms.p,
                                                  c->es.p, gf, c->gm, off, c->axs.p, c->ays.p,
                                                  c->azs.p, c->pots.p, c->cpart.p, c->derr.p, acc,
                                                  c->stream));
        else
          PT_TRY_LAUNCH(c, pt_launch_gravity((int)precision, attempt == 1, c->bk.p, nb,
                                             c->tnode.p, c->gn.p, c->gn32.p, c->xs.p, c->ys.p,
                                             c->zs.p, c->ms.p, c->es.p, gf, off, c->axs.p,
                                             c->ays.p, c->azs.p, c->pots.p, c->cpart.p,
                                             c->derr.p, c->wqueue.p, acc, c->stream));
      }
      if (r + 1 < offs.size())
        PT_TRY_LAUNCH(c, pt_launch_reduce_counters(c->cpart.p, nblk, c->counters.p, c->derr.p,
                                                   c->stream, acc));
      else
        break;
    }
    if (c->cfg.periodic) {
      // LeafEwaldFn after the replicas (examples/Gravity.C:47-50)
      Timer t(c, PT_T_FORCE);
      PT_TRY_HIP(c, ensure(c->ewt, 5 * pt_ewald_table_entries()));
      PT_TRY_HIP(c, ensure(c->ewcount, 1));
      if (hexa)
        PT_TRY_LAUNCH(c, pt_launch_ewald_hexa(c->n, c->xs.p, c->ys.p, c->zs.p, c->nd.p, c->nn,
                                              c->hn.p, c->cfg.period[0], c->cfg.n_replicas,
                                              c->ewt.p, c->ewcount.p, c->axs.p, c->ays.p,
                                              c->azs.p, c->pots.p, c->stream));
      else
        PT_TRY_LAUNCH(c, pt_launch_ewald(c->n, c->xs.p, c->ys.p, c->zs.p, c->nd.p, c->nn,
                                         c->cfg.period[0], c->cfg.n_replicas, c->ewt.p,
                                         c->ewcount.p, c->axs.p, c->ays.p, c->azs.p,
                                         c->pots.p, c->stream));
    }
    // the last replica's counters, the unsort to caller order and the one
    // read-back; a rerun (stack overflow, rare) repeats every replica
    int32_t herr = 0;
    if ((st = finish_counters(c, nblk, &herr, true, offs.size() > 1))) return st;
    if (!herr) break;
    if (attempt == 1) return fail(c, PT_ERR_CAPACITY, "walk stack overflow at 1536 entries");
  }
  if (dual) {
    c->h_counters.n_node_ints += dual_cnt.n_node_ints;
    c->h_counters.n_part_ints += dual_cnt.n_part_ints;
    c->h_counters.n_opens += dual_cnt.n_opens;
    c->h_counters.n_closes += dual_cnt.n_closes;
  }
  if (counters) *counters = c->h_counters;
  c->lists_valid = false;
  c->stage = S_FORCES;
  c->have_forces = true;
  return PT_OK;
}

pt_status pt_compute_forces(pt_ctx* c, double theta, const double offset[3],
                            pt_precision precision, pt_counters* counters) {
  pt_status st;
  if ((st = pt_remake_universe(c, nullptr))) return st;
  if ((st = pt_assign_keys(c))) return st;
  if ((st = sort_and_build(c))) return st;
  if ((st = pt_upward(c))) return st;
  return gravity_fused(c, theta, offset, precision, counters);
}

static pt_status step_particles(pt_ctx* c, double dt, int mode, const char* what) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->have_forces) return fail(c, PT_ERR_STATE, std::string(what) + " before any force evaluation");
  {
    Timer t(c, PT_T_INTEGRATE);
    PT_TRY_LAUNCH(c, pt_launch_kick_drift(c->n, dt, mode, c->x.p, c->y.p, c->z.p, c->vx.p,
                                          c->vy.p, c->vz.p, c->ax.p, c->ay.p, c->az.p,
                                          c->stream));
  }
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  if (mode != PT_STEP_KICK) {
    // positions moved: everything downstream must be rebuilt (Partition::rebuild)
    c->stage = S_PARTICLES;
    c->lists_valid = false;
  }
  return PT_OK;
}

pt_status pt_kick_drift(pt_ctx* c, double dt) {
  return step_particles(c, dt, PT_STEP_KICK_DRIFT, "pt_kick_drift");
}

pt_status pt_kick(pt_ctx* c, double dt) { return step_particles(c, dt, PT_STEP_KICK, "pt_kick"); }

pt_status pt_perturb(pt_ctx* c, double dt) {
  return step_particles(c, dt, PT_STEP_PERTURB, "pt_perturb");
}

pt_status pt_run(pt_ctx* c, int steps, double dt, double theta, pt_precision precision,
                 pt_counters* counters) {
  pt_status st;
  pt_counters total{};
  for (int i = 0; i < steps; i++) {
    pt_counters k{};
    if ((st = pt_compute_forces(c, theta, nullptr, precision, &k))) return st;
    total.n_node_ints += k.n_node_ints;
    total.n_part_ints += k.n_part_ints;
    total.n_opens += k.n_opens;
    total.n_closes += k.n_closes;
    if ((st = pt_kick_drift(c, dt))) return st;
  }
  if (counters) *counters = total;
  return PT_OK;
}

pt_status pt_direct_sum(pt_ctx* c, const int64_t* targets, int64_t nt, double* acc_out) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_PARTICLES, "pt_direct_sum"))) return st;
  if (nt < 0 || (nt > 0 && (!targets || !acc_out))) return fail(c, PT_ERR_ARG, "bad targets");
  for (int64_t i = 0; i < nt; i++)
    if (targets[i] < 0 || targets[i] >= c->n) return fail(c, PT_ERR_ARG, "target index out of range");
  if (nt == 0) return PT_OK;
  int64_t* dt = nullptr;
  double* dout = nullptr;
  PT_TRY_HIP(c, hipMalloc((void**)&dt, nt * sizeof(int64_t)));
  hipError_t e = hipMalloc((void**)&dout, nt * 3 * sizeof(double));
  if (e != hipSuccess) {
    (void)hipFree(dt);
    return hip_fail(c, "hipMalloc", e);
  }
  e = hipMemcpyAsync(dt, targets, nt * sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    if (pt_launch_direct(dt, nt, c->x.p, c->y.p, c->z.p, c->m.p, c->e.p, c->n, dout, c->stream) != 0)
      e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(acc_out, dout, nt * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dt);
  (void)hipFree(dout);
  if (e != hipSuccess) return hip_fail(c, "pt_direct_sum", e);
  return PT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- adjoint

namespace {

// Adjoint work arrays inside c->adj (17n doubles).
struct AdjViews {
  double *lx, *ly, *lz, *lvx, *lvy, *lvz, *lm;   // caller order
  double *wx, *wy, *wz;                          // caller order
  double *wsx, *wsy, *wsz, *gsx, *gsy, *gsz, *gsm;  // sorted order
};

pt_status adj_views(pt_ctx* c, AdjViews* v) {
  const int64_t n = std::max<int64_t>(c->n, 1);
  PT_TRY_HIP(c, ensure(c->adj, 17 * n));
  double* p = c->adj.p;
  double** f[17] = {&v->lx, &v->ly, &v->lz, &v->lvx, &v->lvy, &v->lvz, &v->lm, &v->wx, &v->wy,
                    &v->wz, &v->wsx, &v->wsy, &v->wsz, &v->gsx, &v->gsy, &v->gsz, &v->gsm};
  for (int k = 0; k < 17; k++) *f[k] = p + k * n;
  return PT_OK;
}

pt_status zero(pt_ctx* c, double* p, int64_t count) {
  PT_TRY_LAUNCH(c, pt_launch_fill_u64((unsigned long long*)p, count, 0ull, c->stream));
  return PT_OK;
}

// n x 3 caller array (host or device) -> three device arrays; `stage` is a
// 3n device buffer used for host input.  NULL input gives zeros.
pt_status load3(pt_ctx* c, const double* v, pt_memory where, double* a, double* b, double* d,
                double* stage) {
  const int64_t n = c->n;
  if (!v) {
    pt_status st;
    if ((st = zero(c, a, n)) || (st = zero(c, b, n)) || (st = zero(c, d, n))) return st;
    return PT_OK;
  }
  const double* src = v;
  if (where == PT_HOST) {
    PT_TRY_HIP(c, hipMemcpyAsync(stage, v, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    src = stage;
  }
  PT_TRY_LAUNCH(c, pt_launch_split3(n, src, a, b, d, c->stream));
  return PT_OK;
}

pt_status store3(pt_ctx* c, const double* a, const double* b, const double* d, double* out,
                 pt_memory where, double* stage) {
  const int64_t n = c->n;
  if (!out) return PT_OK;
  double* dst = where == PT_HOST ? stage : out;
  PT_TRY_LAUNCH(c, pt_launch_join3(n, a, b, d, dst, c->stream));
  if (where == PT_HOST) {
    PT_TRY_HIP(c, hipMemcpyAsync(out, stage, (size_t)n * 24, hipMemcpyDeviceToHost, c->stream));
    PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  }
  return PT_OK;
}

pt_status store1(pt_ctx* c, const double* a, double* out, pt_memory where) {
  if (!out) return PT_OK;
  const hipMemcpyKind kind = where == PT_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  PT_TRY_HIP(c, hipMemcpyAsync(out, a, (size_t)c->n * 8, kind, c->stream));
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return PT_OK;
}

pt_status copy_d2d(pt_ctx* c, double* dst, const double* src, int64_t count) {
  PT_TRY_HIP(c, hipMemcpyAsync(dst, src, (size_t)count * 8, hipMemcpyDeviceToDevice, c->stream));
  return PT_OK;
}

pt_status save_state(pt_ctx* c, double* dst) {
  const int64_t n = c->n;
  const double* src[6] = {c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p};
  pt_status st;
  for (int k = 0; k < 6; k++)
    if ((st = copy_d2d(c, dst + k * n, src[k], n))) return st;
  return PT_OK;
}

pt_status load_state(pt_ctx* c, const double* src, bool with_velocities) {
  const int64_t n = c->n;
  double* dst[6] = {c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p};
  pt_status st;
  for (int k = 0; k < (with_velocities ? 6 : 3); k++)
    if ((st = copy_d2d(c, dst[k], src + k * n, n))) return st;
  c->stage = S_PARTICLES;
  c->have_forces = false;
  c->lists_valid = false;
  return PT_OK;
}

// Accelerations at the current positions with the given force model.
pt_status traj_force(pt_ctx* c, int mode, double theta, pt_precision prec, pt_counters* k) {
  if (mode == PT_FORCE_TREE) return pt_compute_forces(c, theta, nullptr, prec, k);
  pt_status st;
  if ((st = need(c, S_PARTICLES, "direct forces"))) return st;
  {
    Timer t(c, PT_T_FORCE);
    PT_TRY_LAUNCH(c, pt_launch_direct_all(c->n, c->x.p, c->y.p, c->z.p, c->m.p, c->e.p, c->ax.p,
                                          c->ay.p, c->az.p, c->pot.p, c->stream));
  }
  if (k) {
    *k = pt_counters{};
    k->n_part_ints = (uint64_t)c->n * (uint64_t)c->n;
  }
  c->have_forces = true;
  return PT_OK;
}

// Groups of consecutive buckets and, for each group, the union of its
// buckets' interaction lists with per-entry member masks (count pass, scan,
// fill pass), for the current tree at opening angle theta.
pt_status group_lists(pt_ctx* c, double theta) {
  pt_status st;
  double gf;
  if ((st = set_visitor(c, theta, nullptr, &gf))) return st;
  const int64_t nb = c->nb;
  const int64_t ns = pt_group_slots(nb);
  if (ns <= 0) return fail(c, PT_ERR_STATE, "no buckets");
  Timer t(c, PT_T_TRAVERSE);
  PT_TRY_HIP(c, ensure(c->gfirst, ns));
  PT_TRY_HIP(c, ensure(c->gnb, ns));
  PT_TRY_HIP(c, ensure(c->gpp_cnt, ns));
  PT_TRY_HIP(c, ensure(c->gpc_cnt, ns));
  PT_TRY_HIP(c, ensure(c->gpp_off, ns + 1));
  PT_TRY_HIP(c, ensure(c->gpc_off, ns + 1));
  if ((st = ensure_scratch(c, pt_scan_scratch_bytes(ns + 1)))) return st;
  PT_TRY_LAUNCH(c, pt_launch_fill_u32((uint32_t*)c->derr.p, 1, 0u, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_group_buckets(c->bk.p, nb, c->gfirst.p, c->gnb.p, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_group_traverse(false, c->tnode.p, c->bk.p, c->gfirst.p, c->gnb.p,
                                            ns, gf, c->offset, c->gpp_cnt.p, c->gpc_cnt.p,
                                            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                            c->derr.p, c->stream));
  PT_TRY_LAUNCH(c, pt_exclusive_scan_i32_to_i64(c->gpp_cnt.p, c->gpp_off.p, ns, c->scratch.p,
                                                c->gpp_off.p + ns, c->stream));
  PT_TRY_LAUNCH(c, pt_exclusive_scan_i32_to_i64(c->gpc_cnt.p, c->gpc_off.p, ns, c->scratch.p,
                                                c->gpc_off.p + ns, c->stream));
  int64_t tot[2] = {0, 0};
  int32_t herr = 0;
  PT_TRY_HIP(c, read_back(c, &tot[0], c->gpp_off.p + ns, 8));
  PT_TRY_HIP(c, read_back(c, &tot[1], c->gpc_off.p + ns, 8));
  PT_TRY_HIP(c, read_back(c, &herr, c->derr.p, 4));
  if (herr) return fail(c, PT_ERR_CAPACITY, "group walk stack overflow");
  c->n_gpp = tot[0];
  c->n_gpc = tot[1];
  PT_TRY_HIP(c, ensure(c->gpp, std::max<int64_t>(tot[0], 1)));
  PT_TRY_HIP(c, ensure(c->gppm, std::max<int64_t>(tot[0], 1)));
  PT_TRY_HIP(c, ensure(c->gpc, std::max<int64_t>(tot[1], 1)));
  PT_TRY_HIP(c, ensure(c->gpcm, std::max<int64_t>(tot[1], 1)));
  PT_TRY_LAUNCH(c, pt_launch_group_traverse(true, c->tnode.p, c->bk.p, c->gfirst.p, c->gnb.p,
                                            ns, gf, c->offset, nullptr, nullptr, c->gpp_off.p,
                                            c->gpc_off.p, c->gpp.p, c->gppm.p, c->gpc.p,
                                            c->gpcm.p, c->derr.p, c->stream));
  return PT_OK;
}

// g (caller order) += (da/dx)^T w, gm += (da/dm)^T w at the current
// positions; w in caller order.
pt_status vjp_accumulate(pt_ctx* c, int mode, double theta, const AdjViews& v, const double* wx,
                         const double* wy, const double* wz, double* gx, double* gy, double* gz,
                         double* gm) {
  pt_status st;
  const int64_t n = c->n;
  if (mode == PT_FORCE_DIRECT) {
    Timer t(c, PT_T_ADJOINT);
    PT_TRY_LAUNCH(c, pt_launch_direct_adjoint(n, c->x.p, c->y.p, c->z.p, c->m.p, c->e.p, wx, wy,
                                              wz, gx, gy, gz, gm, c->stream));
    return PT_OK;
  }
  if (c->adj_prec != PT_FP64 && (st = check_fp32c_range(c))) return st;
  if ((st = pt_remake_universe(c, nullptr))) return st;
  if ((st = pt_assign_keys(c))) return st;
  if ((st = sort_and_build(c))) return st;
  if ((st = pt_upward(c))) return st;
  if ((st = group_lists(c, theta))) return st;
  const int64_t nn = c->nn;
  const int64_t nscr = pt_adjoint_node_scratch(std::max<int64_t>(nn, 1));
  PT_TRY_HIP(c, ensure(c->lam_node, nscr));
  Timer t(c, PT_T_ADJOINT);
  if ((st = zero(c, c->lam_node.p, nscr))) return st;
  if ((st = zero(c, v.gsx, 4 * std::max<int64_t>(n, 1)))) return st;  // gsx gsy gsz gsm
  PT_TRY_LAUNCH(c, pt_launch_gather3(c->perm.p, n, wx, wy, wz, v.wsx, v.wsy, v.wsz, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_adjoint_groups(
                       c->adj_prec, c->bk.p, c->gfirst.p, c->gnb.p, pt_group_slots(c->nb),
                       c->tnode.p, c->gn.p, c->gn32.p, c->xs.p, c->ys.p, c->zs.p, c->ms.p,
                       c->es.p, c->gpp_off.p, c->gpp.p, c->gppm.p, c->gpc_off.p, c->gpc.p,
                       c->gpcm.p, c->offset, v.wsx, v.wsy, v.wsz, v.gsx, v.gsy, v.gsz, v.gsm,
                       c->lam_node.p, nn, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_adjoint_down(n, c->head.p, c->head_pos.p, c->bucket_node.p,
                                          c->nparent.p, c->gn.p, c->lam_node.p, nn, c->xs.p,
                                          c->ys.p, c->zs.p, c->ms.p, c->es.p, v.gsx, v.gsy,
                                          v.gsz, v.gsm, c->stream));
  PT_TRY_LAUNCH(c, pt_launch_scatter_add4(c->perm.p, n, v.gsx, v.gsy, v.gsz, v.gsm, gx, gy, gz,
                                          gm, c->stream));
  return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_set_adjoint_precision(pt_ctx* c, pt_precision precision) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (precision != PT_FP64 && precision != PT_FP32C) return fail(c, PT_ERR_ARG, "bad precision");
  c->adj_prec = precision;
  return PT_OK;
}

pt_status pt_force_vjp(pt_ctx* c, double theta, pt_force_mode mode, const double* w,
                       double* grad_x, double* grad_m, pt_memory where) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_PARTICLES, "pt_force_vjp"))) return st;
  if (mode != PT_FORCE_TREE && mode != PT_FORCE_DIRECT) return fail(c, PT_ERR_ARG, "bad force mode");
  if (!w) return fail(c, PT_ERR_ARG, "missing cotangent");
  if (c->n == 0) return PT_OK;
  if (mode == PT_FORCE_TREE && !(theta > 0.0)) return fail(c, PT_ERR_ARG, "theta must be positive");
  if (mode == PT_FORCE_TREE && c->cfg.hexadecapole)
    return fail(c, PT_ERR_UNSUPPORTED, "no adjoint of the hexadecapole node()");
  if (mode == PT_FORCE_TREE && c->cfg.periodic)
    return fail(c, PT_ERR_UNSUPPORTED, "no adjoint of the periodic replicas");
  AdjViews v;
  if ((st = adj_views(c, &v))) return st;
  // the sorted-order block (gsx..) doubles as the host staging area before
  // and after the tree stages use it
  if ((st = load3(c, w, where, v.wx, v.wy, v.wz, v.gsx))) return st;
  if ((st = zero(c, v.lx, 3 * c->n)) || (st = zero(c, v.lm, c->n))) return st;
  if ((st = vjp_accumulate(c, mode, theta, v, v.wx, v.wy, v.wz, v.lx, v.ly, v.lz, v.lm))) return st;
  if ((st = store3(c, v.lx, v.ly, v.lz, grad_x, where, v.gsx))) return st;
  if ((st = store1(c, v.lm, grad_m, where))) return st;
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return PT_OK;
}

pt_status pt_leapfrog_forward(pt_ctx* c, int32_t steps, double dt, double theta,
                              pt_precision precision, pt_force_mode mode,
                              int32_t checkpoint_every, pt_counters* counters) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_PARTICLES, "pt_leapfrog_forward"))) return st;
  if (steps < 0) return fail(c, PT_ERR_ARG, "negative step count");
  if (mode != PT_FORCE_TREE && mode != PT_FORCE_DIRECT) return fail(c, PT_ERR_ARG, "bad force mode");
  if (precision != PT_FP64 && precision != PT_FP32C) return fail(c, PT_ERR_ARG, "bad precision");
  if (c->n == 0) return fail(c, PT_ERR_ARG, "no particles");
  if (mode == PT_FORCE_TREE && c->cfg.hexadecapole)
    return fail(c, PT_ERR_UNSUPPORTED, "no adjoint of the hexadecapole node()");
  if (mode == PT_FORCE_TREE && c->cfg.periodic)
    return fail(c, PT_ERR_UNSUPPORTED, "no adjoint of the periodic replicas");
  const int32_t every = checkpoint_every > 0 ? checkpoint_every : std::max<int32_t>(steps, 1);
  const int64_t nseg = std::max<int64_t>(1, (steps + every - 1) / every);
  const int64_t n = c->n;
  c->traj_valid = false;
  PT_TRY_HIP(c, ensure(c->ckpt, (nseg + 1) * 6 * n));
  pt_counters total{};
  if (steps == 0 && (st = save_state(c, c->ckpt.p))) return st;
  for (int32_t k = 0; k < steps; k++) {
    if (k % every == 0 && (st = save_state(c, c->ckpt.p + (int64_t)(k / every) * 6 * n))) return st;
    pt_counters kc{};
    if ((st = traj_force(c, mode, theta, precision, &kc))) return st;
    total.n_node_ints += kc.n_node_ints;
    total.n_part_ints += kc.n_part_ints;
    total.n_opens += kc.n_opens;
    total.n_closes += kc.n_closes;
    if ((st = pt_kick_drift(c, dt))) return st;
  }
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  c->traj_valid = true;
  c->traj_steps = steps;
  c->traj_every = every;
  c->traj_mode = mode;
  c->traj_prec = precision;
  c->traj_dt = dt;
  c->traj_theta = theta;
  c->traj_n = n;
  if (counters) *counters = total;
  return PT_OK;
}

pt_status pt_leapfrog_adjoint(pt_ctx* c, const double* lam_x, const double* lam_v,
                              double* grad_x, double* grad_v, double* grad_m, pt_memory where) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->traj_valid || c->traj_n != c->n)
    return fail(c, PT_ERR_STATE, "pt_leapfrog_adjoint without a matching pt_leapfrog_forward");
  const int64_t n = c->n;
  const int32_t steps = c->traj_steps, every = c->traj_every;
  const int64_t nseg = std::max<int64_t>(1, (steps + every - 1) / every);
  const int mode = c->traj_mode;
  const pt_precision prec = (pt_precision)c->traj_prec;
  AdjViews v;
  if ((st = adj_views(c, &v))) return st;
  PT_TRY_HIP(c, ensure(c->seg, (int64_t)std::min(every, std::max(steps, 1)) * 3 * n));
  double* final_state = c->ckpt.p + nseg * 6 * n;
  if ((st = save_state(c, final_state))) return st;
  if ((st = load3(c, lam_x, where, v.lx, v.ly, v.lz, v.gsx))) return st;
  if ((st = load3(c, lam_v, where, v.lvx, v.lvy, v.lvz, v.gsx))) return st;
  if ((st = zero(c, v.lm, n))) return st;
  for (int64_t s = nseg - 1; s >= 0 && steps > 0; s--) {
    const int32_t k0 = (int32_t)(s * every);
    const int32_t k1 = std::min(steps, k0 + every);
    // recompute the segment's positions from its checkpoint
    if ((st = load_state(c, c->ckpt.p + s * 6 * n, true))) return st;
    for (int32_t k = k0; k < k1; k++) {
      double* xk = c->seg.p + (int64_t)(k - k0) * 3 * n;
      if ((st = copy_d2d(c, xk, c->x.p, n)) || (st = copy_d2d(c, xk + n, c->y.p, n)) ||
          (st = copy_d2d(c, xk + 2 * n, c->z.p, n)))
        return st;
      if (k + 1 < k1) {
        if ((st = traj_force(c, mode, c->traj_theta, prec, nullptr))) return st;
        if ((st = pt_kick_drift(c, c->traj_dt))) return st;
      }
    }
    // reverse through the segment
    for (int32_t k = k1 - 1; k >= k0; k--) {
      if ((st = load_state(c, c->seg.p + (int64_t)(k - k0) * 3 * n, false))) return st;
      PT_TRY_LAUNCH(c, pt_launch_adjoint_w(n, c->traj_dt, v.lx, v.ly, v.lz, v.lvx, v.lvy, v.lvz,
                                           v.wx, v.wy, v.wz, c->stream));
      if ((st = vjp_accumulate(c, mode, c->traj_theta, v, v.wx, v.wy, v.wz, v.lx, v.ly, v.lz,
                               v.lm)))
        return st;
    }
  }
  if ((st = store3(c, v.lx, v.ly, v.lz, grad_x, where, v.gsx))) return st;
  if ((st = store3(c, v.lvx, v.lvy, v.lvz, grad_v, where, v.gsx))) return st;
  if ((st = store1(c, v.lm, grad_m, where))) return st;
  if ((st = load_state(c, final_state, true))) return st;
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return PT_OK;
}

}  // extern "C"

extern "C" {

// ---------------------------------------------------------------- read-back

static pt_status d2h(pt_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!dst || bytes == 0) return PT_OK;
  PT_TRY_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return PT_OK;
}

pt_status pt_get_universe(pt_ctx* c, double box[6]) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_UNIVERSE, "pt_get_universe"))) return st;
  return d2h(c, box, c->box.p, 6 * sizeof(double));
}

pt_status pt_get_keys(pt_ctx* c, uint64_t* keys) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_KEYS, "pt_get_keys"))) return st;
  return d2h(c, keys, c->keys.p, (size_t)c->n * 8);
}

pt_status pt_get_sorted(pt_ctx* c, uint64_t* sorted_keys, uint32_t* perm) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_SORTED, "pt_get_sorted"))) return st;
  if ((st = d2h(c, sorted_keys, c->skeys.p, (size_t)c->n * 8))) return st;
  return d2h(c, perm, c->perm.p, (size_t)c->n * 4);
}

pt_status pt_get_tree_size(pt_ctx* c, int64_t* n_nodes, int64_t* n_buckets, int32_t* max_depth) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_TREE, "pt_get_tree_size"))) return st;
  if (n_nodes) *n_nodes = c->nn;
  if (n_buckets) *n_buckets = c->nb;
  if (max_depth) *max_depth = c->max_depth;
  return PT_OK;
}

pt_status pt_get_nodes(pt_ctx* c, uint64_t* key, int32_t* type, int32_t* depth, int32_t* begin,
                       int32_t* count, int32_t* parent, int32_t* first_child) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_TREE, "pt_get_nodes"))) return st;
  const size_t nn = (size_t)c->nn;
  if ((st = d2h(c, key, c->nkey.p, nn * 8))) return st;
  if ((st = d2h(c, type, c->ntype.p, nn * 4))) return st;
  if ((st = d2h(c, depth, c->ndepth.p, nn * 4))) return st;
  if ((st = d2h(c, begin, c->nbegin.p, nn * 4))) return st;
  if ((st = d2h(c, count, c->ncount.p, nn * 4))) return st;
  if ((st = d2h(c, parent, c->nparent.p, nn * 4))) return st;
  return d2h(c, first_child, c->nfirst.p, nn * 4);
}

pt_status pt_get_node_data(pt_ctx* c, double* out) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_MOMENTS, "pt_get_node_data"))) return st;
  const int64_t nn = c->nn;
  std::vector<double> soa((size_t)nn * PT_CD_NFIELDS);
  if ((st = d2h(c, soa.data(), c->nd.p, soa.size() * sizeof(double)))) return st;
  for (int64_t i = 0; i < nn; i++)
    for (int f = 0; f < PT_CD_NFIELDS; f++) out[i * PT_CD_NFIELDS + f] = soa[(size_t)f * nn + i];
  return PT_OK;
}

pt_status pt_get_node_hexadecapole(pt_ctx* c, double* out) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_MOMENTS, "pt_get_node_hexadecapole"))) return st;
  if (!c->cfg.hexadecapole) return fail(c, PT_ERR_STATE, "context not created with hexadecapole");
  const int64_t nn = c->nn;
  std::vector<HexNode> h((size_t)nn);
  if ((st = d2h(c, h.data(), c->hn.p, h.size() * sizeof(HexNode)))) return st;
  for (int64_t i = 0; i < nn; i++) {
    out[i * 22] = h[i].mass;
    for (int k = 0; k < PT_HEXA_MOMS; k++) out[i * 22 + 1 + k] = h[i].h[k];
  }
  return PT_OK;
}

pt_status pt_get_buckets(pt_ctx* c, int32_t* bucket_nodes) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_TREE, "pt_get_buckets"))) return st;
  return d2h(c, bucket_nodes, c->bucket_node.p, (size_t)c->nb * 4);
}

pt_status pt_get_lists_size(pt_ctx* c, int64_t* n_pp, int64_t* n_pc) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->lists_valid) return fail(c, PT_ERR_STATE, "no interaction lists (call pt_traverse)");
  if (n_pp) *n_pp = c->n_pp;
  if (n_pc) *n_pc = c->n_pc;
  return PT_OK;
}

pt_status pt_get_lists(pt_ctx* c, int64_t* pp_off, int32_t* pp, int64_t* pc_off, int32_t* pc) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->lists_valid) return fail(c, PT_ERR_STATE, "no interaction lists (call pt_traverse)");
  const int64_t nb = c->nb;
  std::vector<uint64_t> keys((size_t)c->nn);
  if ((st = d2h(c, keys.data(), c->nkey.p, keys.size() * 8))) return st;
  auto left_aligned = [&](int32_t id) {
    const uint64_t k = keys[(size_t)id];
    const int depth = (63 - __builtin_clzll(k)) / 3;
    return k << (64 - (3 * depth + 1));
  };
  struct L {
    int64_t* off;
    int32_t* ids;
    const DBuf<int64_t>* doff;
    const DBuf<int32_t>* dids;
    int64_t n;
  } lists[2] = {{pp_off, pp, &c->pp_off, &c->pp, c->n_pp}, {pc_off, pc, &c->pc_off, &c->pc, c->n_pc}};
  for (auto& l : lists) {
    if (!l.off || !l.ids) continue;
    if ((st = d2h(c, l.off, l.doff->p, (size_t)(nb + 1) * 8))) return st;
    if ((st = d2h(c, l.ids, l.dids->p, (size_t)l.n * 4))) return st;
    // canonical order inside each bucket: DFS order = ascending left-aligned key
    for (int64_t b = 0; b < nb; b++) {
      std::sort(l.ids + l.off[b], l.ids + l.off[b + 1],
                [&](int32_t a, int32_t q) { return left_aligned(a) < left_aligned(q); });
    }
  }
  return PT_OK;
}

pt_status pt_get_group_lists(pt_ctx* c, double theta, int64_t sizes[3], int32_t* gfirst,
                             int32_t* gnb, int64_t* pp_off, int32_t* pp, uint8_t* ppm,
                             int64_t* pc_off, int32_t* pc, uint8_t* pcm) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_MOMENTS, "pt_get_group_lists"))) return st;
  if (c->cfg.hexadecapole)
    return fail(c, PT_ERR_ARG, "group lists follow the quadrupole open() (no adjoint in the "
                               "HEXADECAPOLE build)");
  if (!sizes) return fail(c, PT_ERR_ARG, "null sizes");
  if ((st = group_lists(c, theta))) return st;
  const int64_t ns = pt_group_slots(c->nb);
  sizes[0] = ns;
  sizes[1] = c->n_gpp;
  sizes[2] = c->n_gpc;
  if (gfirst) {
    if (!gnb || !pp_off || !pp || !ppm || !pc_off || !pc || !pcm)
      return fail(c, PT_ERR_ARG, "null output array");
    const hipMemcpyKind d2h = hipMemcpyDeviceToHost;
    PT_TRY_HIP(c, hipMemcpyAsync(gfirst, c->gfirst.p, ns * 4, d2h, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(gnb, c->gnb.p, ns * 4, d2h, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(pp_off, c->gpp_off.p, (ns + 1) * 8, d2h, c->stream));
    PT_TRY_HIP(c, hipMemcpyAsync(pc_off, c->gpc_off.p, (ns + 1) * 8, d2h, c->stream));
    if (c->n_gpp > 0) {
      PT_TRY_HIP(c, hipMemcpyAsync(pp, c->gpp.p, c->n_gpp * 4, d2h, c->stream));
      PT_TRY_HIP(c, hipMemcpyAsync(ppm, c->gppm.p, c->n_gpp, d2h, c->stream));
    }
    if (c->n_gpc > 0) {
      PT_TRY_HIP(c, hipMemcpyAsync(pc, c->gpc.p, c->n_gpc * 4, d2h, c->stream));
      PT_TRY_HIP(c, hipMemcpyAsync(pcm, c->gpcm.p, c->n_gpc, d2h, c->stream));
    }
    PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  }
  return PT_OK;
}

pt_status pt_get_accelerations(pt_ctx* c, double* acc, double* pot) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (!c->have_forces) return fail(c, PT_ERR_STATE, "pt_get_accelerations before any force evaluation");
  const size_t n = (size_t)c->n;
  std::vector<double> a(n), b(n), d(n);
  if ((st = d2h(c, a.data(), c->ax.p, n * 8))) return st;
  if ((st = d2h(c, b.data(), c->ay.p, n * 8))) return st;
  if ((st = d2h(c, d.data(), c->az.p, n * 8))) return st;
  if (acc) {
    for (size_t i = 0; i < n; i++) {
      acc[3 * i] = a[i];
      acc[3 * i + 1] = b[i];
      acc[3 * i + 2] = d[i];
    }
  }
  return d2h(c, pot, c->pot.p, n * 8);
}

pt_status pt_get_state(pt_ctx* c, double* pos, double* vel) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if ((st = need(c, S_PARTICLES, "pt_get_state"))) return st;
  const size_t n = (size_t)c->n;
  std::vector<double> t(n);
  DBuf<double>* src[2][3] = {{&c->x, &c->y, &c->z}, {&c->vx, &c->vy, &c->vz}};
  double* dst[2] = {pos, vel};
  for (int a = 0; a < 2; a++) {
    if (!dst[a]) continue;
    for (int k = 0; k < 3; k++) {
      if ((st = d2h(c, t.data(), src[a][k]->p, n * 8))) return st;
      for (size_t i = 0; i < n; i++) dst[a][3 * i + k] = t[i];
    }
  }
  return PT_OK;
}

pt_status pt_get_counters(pt_ctx* c, pt_counters* out) {
  if (!c || !out) return PT_ERR_ARG;
  *out = c->h_counters;
  return PT_OK;
}

pt_status pt_set_fused_grouping(pt_ctx* c, int mode) {
  pt_status st = check_ctx(c);
  if (st) return st;
  if (mode < 0 || mode > 64) return fail(c, PT_ERR_ARG, "pt_set_fused_grouping: mode 0..64");
  c->fused_grouping = mode;
  return PT_OK;
}

pt_status pt_set_timing(pt_ctx* c, int enabled) {
  if (!c) return PT_ERR_ARG;
  c->timing = enabled != 0;
  for (int i = 0; i < PT_T_COUNT; i++) c->ev_used[i] = false;
  return PT_OK;
}

pt_status pt_get_timings(pt_ctx* c, float ms[PT_T_COUNT]) {
  pt_status st = check_ctx(c);
  if (st) return st;
  PT_TRY_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < PT_T_COUNT; i++) {
    ms[i] = 0.f;
    if (c->ev_used[i]) (void)hipEventElapsedTime(&ms[i], c->ev[i][0], c->ev[i][1]);
  }
  return PT_OK;
}

}  // extern "C"

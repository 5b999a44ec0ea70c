"""Host-side mirror of ParaTreeT's gravity interface over the gfx950 C ABI.

Names follow the reference so code written against it reads the same:

* ``Configuration``   -- paratreet::Configuration fields and CLI letters
  (src/Configuration.h:84-153) that the gravity path reads.
* ``GravityVisitor``  -- examples/GravityVisitor.h:9-232: constructed from
  (offset, theta); carries gravity_factor / monopole_gravity_factor and the
  CallSelfLeaf / ForceEvenDepth / TargetMustBeLeaf flags.
* ``TreeSpec``        -- src/TreeSpec.h:8-35: tree/decomposition factory; only
  the octree with octree decomposition (one subtree) exists on this path.
* ``Driver``          -- the stages of Driver::run (src/Driver.h:183-288):
  remake_universe, assign_keys, build_tree (sort + recursiveBuild +
  populateTree), start_down(GravityVisitor) (TransposedDownTraverser +
  visitor calls), kick / perturb.

Every call runs on the GPU through libpt_gravity.so; errors raise PtError
with the library's message, as the reference would abort.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib

PT_FP64 = _lib.PT_FP64
PT_FP32C = _lib.PT_FP32C
FORCE_TREE = _lib.PT_FORCE_TREE
FORCE_DIRECT = _lib.PT_FORCE_DIRECT

_TREE_NAMES = {"oct": _lib.PT_TREE_OCT, "binoct": _lib.PT_TREE_BINARY_OCT,
               "kd": _lib.PT_TREE_KD, "longest": _lib.PT_TREE_LONGEST}
_DECOMP_NAMES = {"oct": _lib.PT_DECOMP_OCT, "binoct": _lib.PT_DECOMP_BINARY_OCT,
                 "sfc": _lib.PT_DECOMP_SFC, "kd": _lib.PT_DECOMP_KD,
                 "longest": _lib.PT_DECOMP_LONGEST}


@dataclass
class Configuration:
    """paratreet::Configuration (src/Configuration.h:84-125), gravity subset.
    Defaults follow ExMain::setDefaults (examples/Main.C:26-41) except the
    tree/decomposition types, which default to the octree this path builds."""
    min_n_subtrees: int = 1             # -n
    min_n_partitions: int = 1           # -p
    max_particles_per_leaf: int = 12    # -l
    decomp_type: str = "oct"            # -d
    tree_type: str = "oct"              # -t
    num_iterations: int = 3             # -i
    dSoft: float = 0.0                  # -e
    input_file: str = ""                # -f
    output_file: str = ""               # -v
    periodic: int = 0                   # bPeriodic (parameter file)
    fPeriod: tuple = (0.0, 0.0, 0.0)    # dxPeriod dyPeriod dzPeriod
    nReplicas: int = 0                  # nReplicas
    # the reference's -DHEXADECAPOLE build (examples/Makefile:5): a build flag
    # there, a context option here
    hexadecapole: bool = False
    # the application's dual_tree flag (examples/Main.C:69-71, '-e' there;
    # 'e' is also dSoft's letter, so it is a field here): startDual before
    # startDown in every force evaluation (examples/Gravity.C:30-33)
    dual_tree: bool = False
    origin: dict = field(default_factory=dict)

    LETTERS = {"n": ("min_n_subtrees", int), "p": ("min_n_partitions", int),
               "l": ("max_particles_per_leaf", int), "d": ("decomp_type", str),
               "t": ("tree_type", str), "i": ("num_iterations", int),
               "e": ("dSoft", float), "f": ("input_file", str), "v": ("output_file", str)}
    NAMES = {"nSubtreesMin": "n", "nPartitionsMin": "p", "nParticlesPerLeafMax": "l",
             "achDecompType": "d", "achTreeType": "t", "nIterations": "i", "dSoft": "e",
             "achInputFile": "f", "achOutputFile": "v"}

    def parse(self, argv):
        """Loadable::parse (src/Loadable.C:93-184): '-<letter> value' pairs and
        '-x file' with 'key = value' lines.  Unknown options are returned."""
        rest = []
        i = 0
        while i < len(argv):
            a = argv[i]
            if a == "-x" and i + 1 < len(argv):
                self.load_param_file(argv[i + 1])
                i += 2
                continue
            if len(a) == 2 and a[0] == "-" and a[1] in self.LETTERS and i + 1 < len(argv):
                self._set(a[1], argv[i + 1], "cli")
                i += 2
                continue
            rest.append(a)
            i += 1
        return rest

    def load_param_file(self, path):
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if "=" not in line:
                    continue
                k, v = (s.strip() for s in line.split("=", 1))
                if k in self.NAMES:
                    self._set(self.NAMES[k], v, "file")
                elif k == "bPeriodic":
                    self.periodic = int(v)
                elif k in ("dxPeriod", "dyPeriod", "dzPeriod"):
                    p = list(self.fPeriod)
                    p["xyz".index(k[1])] = float(v)
                    self.fPeriod = tuple(p)
                elif k == "nReplicas":
                    self.nReplicas = int(v)

    def _set(self, letter, value, origin):
        name, typ = self.LETTERS[letter]
        setattr(self, name, typ(value))
        self.origin[name] = origin

    def branch_factor(self):
        return {"oct": 8, "binoct": 2, "kd": 2, "longest": 2}[self.tree_type]

    def to_c(self):
        c = _lib.PtConfig()
        _lib.load().pt_config_default(ctypes.byref(c))
        if self.tree_type not in _TREE_NAMES or self.decomp_type not in _DECOMP_NAMES:
            raise ValueError("invalid tree/decomp type %r/%r" % (self.tree_type, self.decomp_type))
        c.min_n_subtrees = self.min_n_subtrees
        c.min_n_partitions = self.min_n_partitions
        c.max_particles_per_leaf = self.max_particles_per_leaf
        c.decomp_type = _DECOMP_NAMES[self.decomp_type]
        c.tree_type = _TREE_NAMES[self.tree_type]
        c.num_iterations = self.num_iterations
        c.dsoft = self.dSoft
        c.has_dsoft = 1 if "dSoft" in self.origin else 0
        c.periodic = self.periodic
        c.hexadecapole = 1 if self.hexadecapole else 0
        c.dual_tree = 1 if self.dual_tree else 0
        for k in range(3):
            c.period[k] = float(self.fPeriod[k])
        c.n_replicas = int(self.nReplicas)
        return c


class GravityVisitor:
    """examples/GravityVisitor.h:9-20: visitor flags and opening factors."""
    CallSelfLeaf = True
    ForceEvenDepth = True
    TargetMustBeLeaf = True
    opening_geometry_factor_squared = 4.0 / 3.0
    nMinParticleNode = 6

    def __init__(self, offset=(0.0, 0.0, 0.0), theta=0.7):
        self.offset = tuple(float(v) for v in offset)
        self.theta = float(theta)
        self.gravity_factor = self.opening_geometry_factor_squared / (theta * theta)
        self.monopole_gravity_factor = np.sqrt(self.opening_geometry_factor_squared) / theta ** 4


class TreeSpec:
    """src/TreeSpec.h:8-35 restricted to this path: TreeType::eOct with
    DecompType::eOct and a single subtree; other combinations are rejected."""

    def __init__(self, config):
        if config.tree_type != "oct" or config.decomp_type != "oct":
            raise _lib.PtError(_lib.PT_ERR_UNSUPPORTED,
                               "only the octree ('-t oct -d oct') exists on the GPU path")
        self.config = config

    def get_branch_factor(self):
        return 8


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Driver:
    """One GPU context: the particle set and the stages of Driver::run."""

    def __init__(self, config=None, device=0, stream=None):
        self.L = _lib.load()
        self.config = config or Configuration()
        TreeSpec(self.config)
        cfg = self.config.to_c()
        ctx = ctypes.c_void_p()
        st = self.L.pt_ctx_create(ctypes.byref(cfg), int(device), stream, ctypes.byref(ctx))
        if st != 0:
            raise _lib.PtError(st, "pt_ctx_create failed")
        self.ctx = ctx
        self.n = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.L.pt_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != 0:
            raise _lib.PtError(st, self.L.pt_last_error(self.ctx).decode())

    # ---- stages -------------------------------------------------------
    def load(self, pos, mass, soft=None, vel=None):
        """Reader::load + setSoft: original-order host arrays."""
        pos = np.asarray(pos, dtype=np.float64)
        n = len(pos)
        x, y, z = (_f64(pos[:, k]) for k in range(3))
        m = _f64(mass)
        e = None if soft is None else _f64(np.broadcast_to(soft, (n,)))
        if vel is None:
            vx = vy = vz = None
        else:
            vel = np.asarray(vel, dtype=np.float64)
            vx, vy, vz = (_f64(vel[:, k]) for k in range(3))
        self._check(self.L.pt_set_particles(self.ctx, n, _ptr(x), _ptr(y), _ptr(z), _ptr(vx),
                                            _ptr(vy), _ptr(vz), _ptr(m), _ptr(e), _lib.PT_HOST))
        self.n = n

    def load_device(self, n, x, y, z, mass, soft, vx=None, vy=None, vz=None):
        """Same as load() from device pointers (ints), e.g. torch data_ptr()."""
        self._check(self.L.pt_set_particles(self.ctx, int(n), x, y, z, vx, vy, vz, mass, soft,
                                            _lib.PT_DEVICE))
        self.n = int(n)

    def remake_universe(self):
        box = np.zeros(6)
        self._check(self.L.pt_remake_universe(self.ctx, _ptr(box)))
        return box

    def assign_keys(self):
        self._check(self.L.pt_assign_keys(self.ctx))

    def sort(self):
        self._check(self.L.pt_sort(self.ctx))

    def build_tree(self):
        """Subtree::buildTree: sort, recursiveBuild, populateTree."""
        self._check(self.L.pt_sort(self.ctx))
        self._check(self.L.pt_build_tree(self.ctx))
        self._check(self.L.pt_upward(self.ctx))

    def traverse(self, visitor):
        cnt = _lib.PtCounters()
        off = _f64(visitor.offset)
        self._check(self.L.pt_traverse(self.ctx, visitor.theta, _ptr(off), ctypes.byref(cnt)))
        return cnt.as_tuple()

    def interact(self, precision=PT_FP64):
        self._check(self.L.pt_interact(self.ctx, int(precision)))

    def start_dual(self, visitor, source_key=1, target_key=1, precision=PT_FP64):
        """Subtree::startDual<GravityVisitor>: the dual-tree traversal from the
        node pair with these keys (the root pair by default), its lists
        (lists()) and accelerations (accelerations()).  Returns the counters
        (n_node_ints, n_part_ints, n_opens, n_closes), modulo 2^64."""
        cnt = _lib.PtCounters()
        off = _f64(visitor.offset)
        self._check(self.L.pt_start_dual(self.ctx, visitor.theta, _ptr(off), int(source_key),
                                         int(target_key), int(precision), ctypes.byref(cnt)))
        return cnt.as_tuple()

    def start_down(self, visitor, precision=PT_FP64):
        """Partition::startDown<GravityVisitor>: traversal + visitor calls.
        Returns counters (n_node_ints, n_part_ints, n_opens, n_closes)."""
        cnt = self.traverse(visitor)
        self.interact(precision)
        return cnt

    def compute_forces(self, theta=0.7, precision=PT_FP64, offset=(0.0, 0.0, 0.0)):
        cnt = _lib.PtCounters()
        off = _f64(offset)
        self._check(self.L.pt_compute_forces(self.ctx, float(theta), _ptr(off), int(precision),
                                             ctypes.byref(cnt)))
        return cnt.as_tuple()

    def kick_drift(self, dt):
        """Partition::kick then Partition::perturb with timestep dt."""
        self._check(self.L.pt_kick_drift(self.ctx, float(dt)))

    def kick(self, dt):
        """Partition::kick (src/Particle.C:7-9): v += a dt / 2."""
        self._check(self.L.pt_kick(self.ctx, float(dt)))

    def perturb(self, dt):
        """Partition::perturb (src/Particle.C:11-21): v += a dt / 2, x += v dt."""
        self._check(self.L.pt_perturb(self.ctx, float(dt)))

    def run(self, steps, dt, theta=0.7, precision=PT_FP64):
        cnt = _lib.PtCounters()
        self._check(self.L.pt_run(self.ctx, int(steps), float(dt), float(theta), int(precision),
                                  ctypes.byref(cnt)))
        return cnt.as_tuple()

    def direct_sum(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        out = np.zeros((len(t), 3))
        self._check(self.L.pt_direct_sum(self.ctx, _ptr(t), len(t), _ptr(out)))
        return out

    # ---- differentiation (no counterpart in the reference) -----------
    def set_adjoint_precision(self, precision=PT_FP64):
        """Arithmetic of the tree adjoint: PT_FP64 or PT_FP32C (FP32 pairs on
        bucket-centre-shifted coordinates, FP64 sums)."""
        self._check(self.L.pt_set_adjoint_precision(self.ctx, int(precision)))

    def force_vjp(self, w, theta=0.7, mode=FORCE_TREE):
        """(da/dx)^T w and (da/dm)^T w at the current positions (host arrays)."""
        w = _f64(w).reshape(self.n, 3)
        gx = np.zeros((self.n, 3))
        gm = np.zeros(self.n)
        self._check(self.L.pt_force_vjp(self.ctx, float(theta), int(mode), _ptr(w), _ptr(gx),
                                        _ptr(gm), _lib.PT_HOST))
        return gx, gm

    def leapfrog_forward(self, steps, dt, theta=0.7, precision=PT_FP64, mode=FORCE_TREE,
                         checkpoint_every=0):
        """pt_run with checkpoints kept for leapfrog_adjoint; returns counters."""
        cnt = _lib.PtCounters()
        self._check(self.L.pt_leapfrog_forward(self.ctx, int(steps), float(dt), float(theta),
                                               int(precision), int(mode), int(checkpoint_every),
                                               ctypes.byref(cnt)))
        return cnt.as_tuple()

    def leapfrog_adjoint(self, lam_x=None, lam_v=None):
        """Gradients (x0, v0, m) of <lam_x, x_T> + <lam_v, v_T> (host arrays)."""
        lx = None if lam_x is None else _f64(lam_x).reshape(self.n, 3)
        lv = None if lam_v is None else _f64(lam_v).reshape(self.n, 3)
        gx = np.zeros((self.n, 3))
        gv = np.zeros((self.n, 3))
        gm = np.zeros(self.n)
        self._check(self.L.pt_leapfrog_adjoint(self.ctx, _ptr(lx), _ptr(lv), _ptr(gx), _ptr(gv),
                                               _ptr(gm), _lib.PT_HOST))
        return gx, gv, gm

    # ---- read-back ----------------------------------------------------
    def universe(self):
        box = np.zeros(6)
        self._check(self.L.pt_get_universe(self.ctx, _ptr(box)))
        return box

    def keys(self):
        k = np.zeros(self.n, dtype=np.uint64)
        self._check(self.L.pt_get_keys(self.ctx, _ptr(k)))
        return k

    def sorted(self):
        k = np.zeros(self.n, dtype=np.uint64)
        p = np.zeros(self.n, dtype=np.uint32)
        self._check(self.L.pt_get_sorted(self.ctx, _ptr(k), _ptr(p)))
        return k, p

    def tree_size(self):
        nn, nb, md = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._check(self.L.pt_get_tree_size(self.ctx, ctypes.byref(nn), ctypes.byref(nb),
                                            ctypes.byref(md)))
        return nn.value, nb.value, md.value

    def nodes(self):
        nn, _, _ = self.tree_size()
        out = {"key": np.zeros(nn, np.uint64)}
        for f in ("type", "depth", "begin", "count", "parent", "first_child"):
            out[f] = np.zeros(nn, np.int32)
        self._check(self.L.pt_get_nodes(self.ctx, *(_ptr(out[f]) for f in
                                                     ("key", "type", "depth", "begin", "count",
                                                      "parent", "first_child"))))
        return out

    def node_data(self):
        """(n_nodes, 18): radius soft mass cm[3] q[6] lesser[3] greater[3]."""
        nn, _, _ = self.tree_size()
        d = np.zeros((nn, 18))
        self._check(self.L.pt_get_node_data(self.ctx, _ptr(d)))
        return d

    def node_hexadecapole(self):
        """(n_nodes, 22): mass and the reduced moments to fourth order (FMOMR
        order) of a hexadecapole context."""
        nn, _, _ = self.tree_size()
        d = np.zeros((nn, 22))
        self._check(self.L.pt_get_node_hexadecapole(self.ctx, _ptr(d)))
        return d

    def buckets(self):
        _, nb, _ = self.tree_size()
        b = np.zeros(nb, np.int32)
        self._check(self.L.pt_get_buckets(self.ctx, _ptr(b)))
        return b

    def lists(self):
        _, nb, _ = self.tree_size()
        npp, npc = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.L.pt_get_lists_size(self.ctx, ctypes.byref(npp), ctypes.byref(npc)))
        pp_off = np.zeros(nb + 1, np.int64)
        pc_off = np.zeros(nb + 1, np.int64)
        pp = np.zeros(max(1, npp.value), np.int32)
        pc = np.zeros(max(1, npc.value), np.int32)
        self._check(self.L.pt_get_lists(self.ctx, _ptr(pp_off), _ptr(pp), _ptr(pc_off), _ptr(pc)))
        return pp_off, pp[:npp.value], pc_off, pc[:npc.value]

    def group_lists(self, theta=0.7):
        """Bucket groups and member-masked lists of the tree adjoint for the
        current tree: (gfirst, gnb, pp_off, pp, ppm, pc_off, pc, pcm)."""
        sizes = np.zeros(3, np.int64)
        self._check(self.L.pt_get_group_lists(self.ctx, float(theta), _ptr(sizes), *([None] * 8)))
        ns, npp, npc = (int(v) for v in sizes)
        gfirst = np.zeros(ns, np.int32)
        gnb = np.zeros(ns, np.int32)
        pp_off = np.zeros(ns + 1, np.int64)
        pc_off = np.zeros(ns + 1, np.int64)
        pp = np.zeros(max(1, npp), np.int32)
        pc = np.zeros(max(1, npc), np.int32)
        ppm = np.zeros(max(1, npp), np.uint8)
        pcm = np.zeros(max(1, npc), np.uint8)
        self._check(self.L.pt_get_group_lists(self.ctx, float(theta), _ptr(sizes), _ptr(gfirst),
                                              _ptr(gnb), _ptr(pp_off), _ptr(pp), _ptr(ppm),
                                              _ptr(pc_off), _ptr(pc), _ptr(pcm)))
        return gfirst, gnb, pp_off, pp[:npp], ppm[:npp], pc_off, pc[:npc], pcm[:npc]

    def accelerations(self):
        acc = np.zeros((self.n, 3))
        pot = np.zeros(self.n)
        self._check(self.L.pt_get_accelerations(self.ctx, _ptr(acc), _ptr(pot)))
        return acc, pot

    def state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        self._check(self.L.pt_get_state(self.ctx, _ptr(pos), _ptr(vel)))
        return pos, vel

    def set_fused_grouping(self, mode=True):
        """The walk of compute_forces: True / 1 automatic (the default),
        False / 0 one walk per bucket, 2..64 bucket groups of at most that
        many particles; the same calls and counters in every mode."""
        self._check(self.L.pt_set_fused_grouping(self.ctx, int(mode)))

    def set_timing(self, on=True):
        self._check(self.L.pt_set_timing(self.ctx, 1 if on else 0))

    def timings(self):
        ms = (ctypes.c_float * len(_lib.TIMERS))()
        self._check(self.L.pt_get_timings(self.ctx, ms))
        return dict(zip(_lib.TIMERS, list(ms)))

    def device_view(self):
        v = _lib.PtDeviceView()
        self._check(self.L.pt_get_device_view(self.ctx, ctypes.byref(v)))
        return v

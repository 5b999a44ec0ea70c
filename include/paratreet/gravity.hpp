// paratreet/gravity.hpp -- C++ host interface of the MI355X gravity path.
//
// Mirrors the reference's user-facing types for this path so code written
// against them reads the same:
//   paratreet::Configuration  -- src/Configuration.h:84-187 (fields, CLI
//                                letters, '-x' key = value files as
//                                Loadable::parse, src/Loadable.C:93-184)
//   paratreet::TreeSpec       -- src/TreeSpec.h:8-35 (octree only here)
//   paratreet::GravityVisitor -- examples/GravityVisitor.h:9-20 (flags and
//                                opening factors)
//   paratreet::CentroidData   -- examples/CentroidData.h:24-91 (node data as
//                                computed on the GPU)
//   paratreet::Driver         -- the Driver::run stages (src/Driver.h:183-288)
//                                with Partition::startDown<GravityVisitor>
//                                (src/Partition.h:129-134)
// Everything computes on the GPU through the C ABI of pt_gravity.h; errors
// throw paratreet::Error carrying the library's status and message.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../pt_gravity.h"

namespace paratreet {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

enum class TreeType { eOct = 0, eBinaryOct, eKd, eLongest, eInvalid = 100 };
enum class DecompType { eOct = 0, eBinaryOct, eSfc, eKd, eLongest, eInvalid = 100 };

inline TreeType treeTypeFromString(const std::string& s) {
  if (s == "oct") return TreeType::eOct;
  if (s == "binoct") return TreeType::eBinaryOct;
  if (s == "kd") return TreeType::eKd;
  if (s == "longest") return TreeType::eLongest;
  throw Error(PT_ERR_ARG, "-- invalid tree type value, " + s + " --");
}

inline DecompType decompTypeFromString(const std::string& s) {
  if (s == "oct") return DecompType::eOct;
  if (s == "binoct") return DecompType::eBinaryOct;
  if (s == "sfc") return DecompType::eSfc;
  if (s == "kd") return DecompType::eKd;
  if (s == "longest") return DecompType::eLongest;
  throw Error(PT_ERR_ARG, "-- invalid decomp type value, " + s + " --");
}

// paratreet::Configuration, gravity subset.  Defaults follow
// ExMain::setDefaults (examples/Main.C:26-41) except tree/decomposition,
// which default to the octree this path implements.
struct Configuration {
  int min_n_subtrees = 1;          // -n nSubtreesMin
  int min_n_partitions = 1;        // -p nPartitionsMin
  int max_particles_per_leaf = 12; // -l nParticlesPerLeafMax
  DecompType decomp_type = DecompType::eOct;  // -d achDecompType
  TreeType tree_type = TreeType::eOct;        // -t achTreeType
  int num_iterations = 3;          // -i nIterations
  double dSoft = 0.0;              // -e dSoft
  bool dSoft_set = false;          // Configuration::origin_of("dSoft") != Unknown
  std::string input_file;          // -f achInputFile
  std::string output_file;         // -v achOutputFile
  int periodic = 0;                // bPeriodic
  double fPeriod[3] = {0, 0, 0};   // dxPeriod dyPeriod dzPeriod
  int nReplicas = 0;               // nReplicas
  bool hexadecapole = false;       // the -DHEXADECAPOLE build (examples/Makefile:5)
  bool dual_tree = false;          // the application's dual_tree ('-e', examples/Main.C:69-71)

  int branchFactor() const {
    switch (tree_type) {
      case TreeType::eOct: return 8;
      case TreeType::eBinaryOct:
      case TreeType::eKd:
      case TreeType::eLongest: return 2;
      default: throw Error(PT_ERR_ARG, "Tree type is invalid");
    }
  }

  // Loadable::parse: "-<letter> value" pairs and "-x file"; returns the
  // arguments it did not consume (application flags such as -m).
  std::vector<std::string> parse(int argc, char** argv) {
    std::vector<std::string> rest;
    for (int i = 1; i < argc; i++) {
      const std::string a = argv[i];
      if (a.size() == 2 && a[0] == '-' && i + 1 < argc && setLetter(a[1], argv[i + 1])) {
        i++;
      } else if (a == "-x" && i + 1 < argc) {
        loadParamFile(argv[++i]);
      } else {
        rest.push_back(a);
      }
    }
    return rest;
  }

  void loadParamFile(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw Error(PT_ERR_ARG, "cannot open parameter file " + path);
    std::string line;
    while (std::getline(f, line)) {
      const auto hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
      const auto eq = line.find('=');
      if (eq == std::string::npos) continue;
      const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
      static const char* names[][2] = {{"nSubtreesMin", "n"}, {"nPartitionsMin", "p"},
                                       {"nParticlesPerLeafMax", "l"}, {"achDecompType", "d"},
                                       {"achTreeType", "t"}, {"nIterations", "i"},
                                       {"dSoft", "e"}, {"achInputFile", "f"},
                                       {"achOutputFile", "v"}};
      for (auto& nm : names)
        if (k == nm[0]) setLetter(nm[1][0], v);
      // fields registered without a command-line letter (src/Configuration.h:145-149)
      if (k == "bPeriodic") periodic = std::atoi(v.c_str());
      if (k == "dxPeriod") fPeriod[0] = std::atof(v.c_str());
      if (k == "dyPeriod") fPeriod[1] = std::atof(v.c_str());
      if (k == "dzPeriod") fPeriod[2] = std::atof(v.c_str());
      if (k == "nReplicas") nReplicas = std::atoi(v.c_str());
    }
  }

  pt_config toC() const {
    pt_config c;
    pt_config_default(&c);
    c.min_n_subtrees = min_n_subtrees;
    c.min_n_partitions = min_n_partitions;
    c.max_particles_per_leaf = max_particles_per_leaf;
    c.decomp_type = (int32_t)decomp_type;
    c.tree_type = (int32_t)tree_type;
    c.num_iterations = num_iterations;
    c.dsoft = dSoft;
    c.has_dsoft = dSoft_set ? 1 : 0;
    c.periodic = periodic;
    c.hexadecapole = hexadecapole ? 1 : 0;
    for (int k = 0; k < 3; k++) c.period[k] = fPeriod[k];
    c.n_replicas = nReplicas;
    c.dual_tree = dual_tree ? 1 : 0;
    return c;
  }

 private:
  static std::string trim(const std::string& s) {
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
  }
  bool setLetter(char c, const std::string& v) {
    switch (c) {
      case 'n': min_n_subtrees = std::atoi(v.c_str()); return true;
      case 'p': min_n_partitions = std::atoi(v.c_str()); return true;
      case 'l': max_particles_per_leaf = std::atoi(v.c_str()); return true;
      case 'd': decomp_type = decompTypeFromString(v); return true;
      case 't': tree_type = treeTypeFromString(v); return true;
      case 'i': num_iterations = std::atoi(v.c_str()); return true;
      case 'e': dSoft = std::atof(v.c_str()); dSoft_set = true; return true;
      case 'f': input_file = v; return true;
      case 'v': output_file = v; return true;
      default: return false;
    }
  }
};

// TreeSpec restricted to this path: TreeType::eOct with DecompType::eOct and
// one subtree per GPU.  Other combinations are rejected with an error.
class TreeSpec {
 public:
  explicit TreeSpec(const Configuration& c) : config_(c) {
    if (c.tree_type != TreeType::eOct || c.decomp_type != DecompType::eOct)
      throw Error(PT_ERR_UNSUPPORTED, "only the octree ('-t oct -d oct') runs on the GPU path");
  }
  int getBranchFactor() const { return 8; }

 private:
  Configuration config_;
};

// GravityVisitor (examples/GravityVisitor.h:9-20): constructed exactly as the
// reference does; the open/node/leaf bodies run on the GPU.
class GravityVisitor {
 public:
  static constexpr bool CallSelfLeaf = true;
  static constexpr bool ForceEvenDepth = true;
  static constexpr bool TargetMustBeLeaf = true;
  static constexpr double opening_geometry_factor_squared = 4.0 / 3.0;
  static constexpr int nMinParticleNode = 6;

  GravityVisitor() : offset{0, 0, 0}, theta(0.7) { setFactors(); }
  GravityVisitor(const double off[3], double th) : offset{off[0], off[1], off[2]}, theta(th) {
    setFactors();
  }

  double offset[3];
  double theta;
  double gravity_factor = 0;
  double monopole_gravity_factor = 0;

 private:
  void setFactors() {
    gravity_factor = opening_geometry_factor_squared / (theta * theta);
    monopole_gravity_factor = std::sqrt(opening_geometry_factor_squared) / (theta * theta * theta * theta);
  }
};

// CentroidData as computed on the GPU (MultipoleMoments + OrientedBox + count).
struct CentroidData {
  double radius, soft, totalMass;
  double cm[3];
  double xx, xy, xz, yy, yz, zz;
  double lesser[3], greater[3];
  int32_t count;
};

// Tipsy standard binary files (big-endian), as Reader::load consumes them
// (src/Reader.C:11-108) and TipsyWriter writes them (src/Writer.C:87-159):
// a 32-byte header {double time; int nbodies, ndim, nsph, ndark, nstar, pad},
// then gas, dark and star records of mass, pos[3], vel[3] and the per-type
// floats (gas: rho temp hsmooth metals phi; dark: eps phi; star: metals tform
// eps phi).  Positions and velocities are float (48 / 36 / 44-byte records)
// or double (72 / 60 / 68 bytes, TipsyWriter's use_double); the reader tells
// them apart by the file size.  soft = gas hsmooth, dark eps, 0 for stars as
// Reader.C sets them (Reader::setSoft overrides it when dSoft is configured).
enum TipsyType : int8_t { kTipsyStar = 1, kTipsyGas = 2, kTipsyDark = 3 };  // Particle::Type

struct TipsyData {
  double time = 0;
  std::vector<double> pos, vel, mass, soft;  // pos, vel: n x 3
  std::vector<int8_t> type;                  // TipsyType per particle, file order
  bool double_records = false;
};

namespace tipsy_detail {
inline uint64_t be_load(const unsigned char* b, int bytes) {
  uint64_t v = 0;
  for (int i = 0; i < bytes; i++) v = (v << 8) | b[i];
  return v;
}
inline void be_store(unsigned char* b, uint64_t v, int bytes) {
  for (int i = bytes - 1; i >= 0; i--, v >>= 8) b[i] = (unsigned char)(v & 0xff);
}
inline double be_f32(const unsigned char* b) {
  const uint32_t u = (uint32_t)be_load(b, 4);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline double be_f64(const unsigned char* b) {
  const uint64_t u = be_load(b, 8);
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
inline void put_f32(unsigned char* b, double v) {
  const float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  be_store(b, u, 4);
}
inline void put_f64(unsigned char* b, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  be_store(b, u, 8);
}
// floats after vel: gas 5, dark 2, star 4
inline int tail_floats(int kind) { return kind == 0 ? 5 : kind == 1 ? 2 : 4; }
inline size_t record_bytes(int kind, bool dbl) {
  return 4 + (dbl ? 48 : 24) + 4 * (size_t)tail_floats(kind);
}
}  // namespace tipsy_detail

inline TipsyData readTipsy(const std::string& path) {
  using namespace tipsy_detail;
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw Error(PT_ERR_ARG, "cannot open tipsy file " + path);
  const uint64_t size = (uint64_t)f.tellg();
  f.seekg(0);
  auto rd = [&](void* p, size_t n) {
    f.read((char*)p, (std::streamsize)n);
    if (!f) throw Error(PT_ERR_ARG, "short tipsy file " + path);
  };
  unsigned char h[32];
  rd(h, 32);
  TipsyData t;
  t.time = be_f64(h);
  const int64_t nbodies = (int32_t)be_load(h + 8, 4), nsph = (int32_t)be_load(h + 16, 4),
                ndark = (int32_t)be_load(h + 20, 4), nstar = (int32_t)be_load(h + 24, 4);
  if (nsph < 0 || ndark < 0 || nstar < 0 || nsph + ndark + nstar != nbodies)
    throw Error(PT_ERR_ARG, "inconsistent tipsy header in " + path);
  const int64_t cnt[3] = {nsph, ndark, nstar};
  int layout = -1;
  for (int dbl = 0; dbl < 2 && layout < 0; dbl++) {
    uint64_t body = 0;
    for (int k = 0; k < 3; k++) body += (uint64_t)cnt[k] * record_bytes(k, dbl != 0);
    if (size == 32 + body) layout = dbl;
  }
  if (layout < 0) throw Error(PT_ERR_ARG, "tipsy file size matches neither float nor double records: " + path);
  t.double_records = layout == 1;
  const int vb = t.double_records ? 8 : 4;
  static const int8_t types[3] = {kTipsyGas, kTipsyDark, kTipsyStar};
  // position of the softening among the trailing floats (-1: none, stars)
  static const int soft_at[3] = {2, 0, -1};
  t.pos.reserve(3 * nbodies);
  t.vel.reserve(3 * nbodies);
  for (int k = 0; k < 3; k++) {
    const size_t rb = record_bytes(k, t.double_records);
    std::vector<unsigned char> raw(rb * (size_t)cnt[k]);
    if (!raw.empty()) rd(raw.data(), raw.size());
    for (int64_t i = 0; i < cnt[k]; i++) {
      const unsigned char* r = &raw[rb * (size_t)i];
      t.mass.push_back(be_f32(r));
      for (int d = 0; d < 3; d++) {
        t.pos.push_back(vb == 8 ? be_f64(r + 4 + 8 * d) : be_f32(r + 4 + 4 * d));
        t.vel.push_back(vb == 8 ? be_f64(r + 4 + 3 * vb + 8 * d) : be_f32(r + 4 + 3 * vb + 4 * d));
      }
      const unsigned char* tail = r + 4 + 6 * vb;
      t.soft.push_back(soft_at[k] >= 0 ? be_f32(tail + 4 * soft_at[k]) : 0.0);
      t.type.push_back(types[k]);
    }
  }
  return t;
}

// TipsyWriter::do_write (src/Writer.C:100-159): particles in Particle::order
// (here: their index, the order the reader assigned), each as the record of
// its type with mass, position and velocity set and the other fields zero,
// positions and velocities in double when use_double (the reference's choice
// when Real is double).  The types must run gas, then dark, then star.
inline void writeTipsy(const std::string& path, double time, const std::vector<double>& pos,
                       const std::vector<double>& vel, const std::vector<double>& mass,
                       const std::vector<int8_t>& type, bool use_double = true) {
  using namespace tipsy_detail;
  const int64_t n = (int64_t)mass.size();
  if ((int64_t)type.size() != n || (int64_t)pos.size() != 3 * n || (int64_t)vel.size() != 3 * n)
    throw Error(PT_ERR_ARG, "writeTipsy: array sizes differ");
  int64_t cnt[3] = {0, 0, 0};
  int last = 0;
  for (int64_t i = 0; i < n; i++) {
    const int k = type[i] == kTipsyGas ? 0 : type[i] == kTipsyDark ? 1 : type[i] == kTipsyStar ? 2 : -1;
    if (k < 0) throw Error(PT_ERR_ARG, "writeTipsy: particle type outside gas/dark/star");
    if (k < last) throw Error(PT_ERR_ARG, "writeTipsy: records must run gas, dark, star");
    last = k;
    cnt[k]++;
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw Error(PT_ERR_ARG, "cannot write " + path);
  unsigned char h[32] = {0};
  put_f64(h, time);
  be_store(h + 8, (uint64_t)n, 4);
  be_store(h + 12, 3, 4);
  for (int k = 0; k < 3; k++) be_store(h + 16 + 4 * k, (uint64_t)cnt[k], 4);
  bool ok = std::fwrite(h, 1, 32, f) == 32;
  std::vector<unsigned char> r;
  for (int64_t i = 0; i < n && ok; i++) {
    const int k = type[i] == kTipsyGas ? 0 : type[i] == kTipsyDark ? 1 : 2;
    r.assign(record_bytes(k, use_double), 0);
    put_f32(r.data(), mass[i]);
    for (int d = 0; d < 3; d++) {
      if (use_double) {
        put_f64(r.data() + 4 + 8 * d, pos[3 * i + d]);
        put_f64(r.data() + 28 + 8 * d, vel[3 * i + d]);
      } else {
        put_f32(r.data() + 4 + 4 * d, pos[3 * i + d]);
        put_f32(r.data() + 16 + 4 * d, vel[3 * i + d]);
      }
    }
    ok = std::fwrite(r.data(), 1, r.size(), f) == r.size();
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) throw Error(PT_ERR_ARG, "short write to " + path);
}

// TipsyWriter's file name (src/Writer.C:112).
inline std::string tipsySnapshotName(const std::string& prefix, int iteration) {
  return prefix + "." + std::to_string(iteration) + ".tipsy";
}

// The stages of one Driver::run iteration for the gravity visitor on one GPU.
class Driver {
 public:
  explicit Driver(const Configuration& cfg, int device = 0) : config_(cfg), spec_(cfg) {
    const pt_config c = cfg.toC();
    check(pt_ctx_create(&c, device, nullptr, &ctx_), "pt_ctx_create");
  }
  ~Driver() { pt_ctx_destroy(ctx_); }
  Driver(const Driver&) = delete;
  Driver& operator=(const Driver&) = delete;

  // Reader::load + setSoft: original-order host arrays.
  void load(const std::vector<double>& pos, const std::vector<double>& vel,
            const std::vector<double>& mass, const std::vector<double>& soft) {
    n_ = (int64_t)mass.size();
    std::vector<double> x(n_), y(n_), z(n_), vx(n_), vy(n_), vz(n_);
    for (int64_t i = 0; i < n_; i++) {
      x[i] = pos[3 * i]; y[i] = pos[3 * i + 1]; z[i] = pos[3 * i + 2];
      vx[i] = vel.empty() ? 0.0 : vel[3 * i];
      vy[i] = vel.empty() ? 0.0 : vel[3 * i + 1];
      vz[i] = vel.empty() ? 0.0 : vel[3 * i + 2];
    }
    check(pt_set_particles(ctx_, n_, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                           mass.data(), soft.empty() ? nullptr : soft.data(), PT_HOST),
          "pt_set_particles");
    mass_ = mass;
  }

  // remakeUniverse + assignKeys + Subtree::buildTree (sort, recursiveBuild,
  // populateTree).
  void buildTree() {
    double box[6];
    check(pt_remake_universe(ctx_, box), "pt_remake_universe");
    check(pt_assign_keys(ctx_), "pt_assign_keys");
    check(pt_sort(ctx_), "pt_sort");
    check(pt_build_tree(ctx_), "pt_build_tree");
    check(pt_upward(ctx_), "pt_upward");
  }

  // Partition::startDown<GravityVisitor>: interaction lists + visitor calls.
  pt_counters startDown(const GravityVisitor& v, pt_precision prec = PT_FP64) {
    pt_counters c{};
    check(pt_traverse(ctx_, v.theta, v.offset, &c), "pt_traverse");
    check(pt_interact(ctx_, prec), "pt_interact");
    return c;
  }

  // Subtree::startDual<GravityVisitor>: the dual-tree traversal from the node
  // pair with these keys (1, 1: the root pair), evaluated in `prec`.
  pt_counters startDual(const GravityVisitor& v, uint64_t source_key = 1, uint64_t target_key = 1,
                        pt_precision prec = PT_FP64) {
    pt_counters c{};
    check(pt_start_dual(ctx_, v.theta, v.offset, source_key, target_key, prec, &c),
          "pt_start_dual");
    return c;
  }

  // Everything above in one fused evaluation (no lists kept).
  pt_counters computeForces(const GravityVisitor& v, pt_precision prec = PT_FP64) {
    pt_counters c{};
    check(pt_compute_forces(ctx_, v.theta, v.offset, prec, &c), "pt_compute_forces");
    return c;
  }

  // Partition::kick + Partition::perturb.
  void kickDrift(double dt) { check(pt_kick_drift(ctx_, dt), "pt_kick_drift"); }
  // Partition::kick (v += a dt / 2) and Partition::perturb (v += a dt / 2,
  // x += v dt) apart, with the iteration's output between them.
  void kick(double dt) { check(pt_kick(ctx_, dt), "pt_kick"); }
  void perturb(double dt) { check(pt_perturb(ctx_, dt), "pt_perturb"); }

  std::vector<double> accelerations() {
    std::vector<double> acc(3 * n_), pot(n_);
    check(pt_get_accelerations(ctx_, acc.data(), pot.data()), "pt_get_accelerations");
    return acc;
  }

  std::vector<CentroidData> nodeData() {
    int64_t nn = 0, nb = 0;
    int32_t md = 0;
    check(pt_get_tree_size(ctx_, &nn, &nb, &md), "pt_get_tree_size");
    std::vector<double> raw(18 * nn);
    std::vector<int32_t> count(nn);
    check(pt_get_node_data(ctx_, raw.data()), "pt_get_node_data");
    check(pt_get_nodes(ctx_, nullptr, nullptr, nullptr, nullptr, count.data(), nullptr, nullptr),
          "pt_get_nodes");
    std::vector<CentroidData> out(nn);
    for (int64_t i = 0; i < nn; i++) {
      const double* r = &raw[18 * i];
      CentroidData& d = out[i];
      d.radius = r[0]; d.soft = r[1]; d.totalMass = r[2];
      for (int k = 0; k < 3; k++) { d.cm[k] = r[3 + k]; d.lesser[k] = r[12 + k]; d.greater[k] = r[15 + k]; }
      d.xx = r[6]; d.xy = r[7]; d.xz = r[8]; d.yy = r[9]; d.yz = r[10]; d.zz = r[11];
      d.count = count[i];
    }
    return out;
  }

  // Writer::do_write format (src/Writer.C:31-72): N, then the x, y and z
  // blocks, one "%.14g" value per line, original particle order.
  void outputParticleAccelerations(const std::string& prefix) {
    const std::vector<double> acc = accelerations();
    FILE* f = std::fopen((prefix + ".acc").c_str(), "w");
    if (!f) throw Error(PT_ERR_ARG, "cannot write " + prefix + ".acc");
    std::fprintf(f, "%lld\n", (long long)n_);
    for (int k = 0; k < 3; k++)
      for (int64_t i = 0; i < n_; i++) std::fprintf(f, "%.14g\n", acc[3 * i + k]);
    std::fclose(f);
  }

  // paratreet::outputTipsy (src/Paratreet.h:252-259) -> TipsyWriter: the
  // current positions and velocities with the loaded masses, in particle
  // order, to <prefix>.<iteration>.tipsy; returns the file name.  `type`
  // gives each particle's Tipsy type (empty: all dark).
  std::string outputTipsy(const std::string& prefix, int iteration, double time,
                          const std::vector<int8_t>& type = {}, bool use_double = true) {
    std::vector<double> pos(3 * n_), vel(3 * n_);
    check(pt_get_state(ctx_, pos.data(), vel.data()), "pt_get_state");
    const std::string name = tipsySnapshotName(prefix, iteration);
    writeTipsy(name, time, pos, vel, mass_,
               type.empty() ? std::vector<int8_t>((size_t)n_, (int8_t)kTipsyDark) : type, use_double);
    return name;
  }

  int64_t size() const { return n_; }
  pt_ctx* context() { return ctx_; }
  const Configuration& config() const { return config_; }

 private:
  void check(int st, const char* what) {
    if (st != PT_OK) throw Error(st, std::string(what) + ": " + (ctx_ ? pt_last_error(ctx_) : "no context"));
  }
  Configuration config_;
  TreeSpec spec_;
  pt_ctx* ctx_ = nullptr;
  int64_t n_ = 0;
  std::vector<double> mass_;
};

}  // namespace paratreet

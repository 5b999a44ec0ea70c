// Gravity application on the MI355X path: examples/Main.C + examples/Gravity.C
// restricted to the non-periodic GravityVisitor traversal on one GPU.
//
//   pt_gravity -f <tipsy> [-l B] [-e dSoft] [-i iterations] [-v out_prefix]
//              [-m] [-j max_timestep] [-theta t] [-fp32c] [-hexadecapole]
//              [-snapshot K] [-dual] [-t oct -d oct]
//
// Per iteration (Driver::run, src/Driver.h:183-288): build the tree, run
// startDown<GravityVisitor>, kick, then postIterationFn -- the accelerations
// at iteration 0 when -v is given (examples/Gravity.C:54-61) and, with
// -snapshot K, a Tipsy snapshot <out>.<iter>.tipsy every K iterations (the
// paratreet::outputTipsy call of examples/Collision.C:64) -- then perturb,
// with timestep min(L / v_max / N^(1/3), max_timestep)
// (examples/Gravity.C:63-67).  The snapshot time is the file's time plus the
// timesteps of the earlier iterations (Partition::time_advanced).  -dual is
// the reference application's dual_tree flag ('-e' in examples/Main.C:69-71,
// where Configuration also registers 'e' for dSoft): Subtree::startDual
// before startDown in every iteration (examples/Gravity.C:30-33).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "paratreet/gravity.hpp"

int main(int argc, char** argv) {
  using namespace paratreet;
  try {
    Configuration conf;
    conf.max_particles_per_leaf = 12;  // examples/Main.C:29
    std::vector<std::string> rest = conf.parse(argc, argv);
    double theta = 0.7;           // examples/Main.C:55
    double max_timestep = 1e-5;   // examples/Main.C:57
    pt_precision prec = PT_FP64;
    int snapshot_every = 0;
    for (size_t i = 0; i < rest.size(); i++) {
      if (rest[i] == "-m") continue;  // Morton keys: the only key of this path
      if (rest[i] == "-j" && i + 1 < rest.size()) max_timestep = std::atof(rest[++i].c_str());
      else if (rest[i] == "-theta" && i + 1 < rest.size()) theta = std::atof(rest[++i].c_str());
      else if (rest[i] == "-fp32c") prec = PT_FP32C;
      else if (rest[i] == "-hexadecapole") conf.hexadecapole = true;  // -DHEXADECAPOLE build
      else if (rest[i] == "-dual") conf.dual_tree = true;
      else if (rest[i] == "-snapshot" && i + 1 < rest.size())
        snapshot_every = std::atoi(rest[++i].c_str());
      else {
        std::fprintf(stderr, "unsupported option %s\n", rest[i].c_str());
        return 2;
      }
    }
    if (conf.input_file.empty()) {
      std::fprintf(stderr, "usage: %s -f <tipsy> [-l B] [-e dSoft] [-i it] [-v out] [-theta t] [-fp32c] [-hexadecapole] [-snapshot K] [-dual]\n",
                   argv[0]);
      return 2;
    }
    TipsyData t = readTipsy(conf.input_file);
    Driver d(conf);
    d.load(t.pos, t.vel, t.mass, t.soft);
    const int64_t n = d.size();
    const double zero[3] = {0, 0, 0};
    double time_advanced = t.time;
    for (int iter = 0; iter < conf.num_iterations; iter++) {
      const pt_counters c = d.computeForces(GravityVisitor(zero, theta), prec);
      std::printf("* Iteration %d: %llu node-particle interactions, %llu bucket-particle interactions "
                  "%llu node opens, %llu node closes\n",
                  iter, (unsigned long long)c.n_node_ints, (unsigned long long)c.n_part_ints,
                  (unsigned long long)c.n_opens, (unsigned long long)c.n_closes);
      // getTimestep: universe side / v_max / N^(1/3), capped at max_timestep
      std::vector<double> pos(3 * n), vel(3 * n);
      pt_get_state(d.context(), pos.data(), vel.data());
      double vmax = 0, box[6];
      pt_get_universe(d.context(), box);
      for (int64_t i = 0; i < n; i++)
        vmax = std::max(vmax, std::sqrt(vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                                        vel[3 * i + 2] * vel[3 * i + 2]));
      const double len = box[3] - box[0];
      const double dt = vmax > 0 ? std::min(len / vmax / std::cbrt((double)n), max_timestep)
                                 : max_timestep;
      d.kick(dt);
      if (iter == 0 && !conf.output_file.empty()) d.outputParticleAccelerations(conf.output_file);
      if (snapshot_every > 0 && iter % snapshot_every == 0 && !conf.output_file.empty())
        std::printf("Outputting to Tipsy file %s\n",
                    d.outputTipsy(conf.output_file, iter, time_advanced, t.type).c_str());
      d.perturb(dt);
      time_advanced += dt;
    }
    return 0;
  } catch (const paratreet::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
    return 1;
  }
}

// Gravity application on the MI355X path: examples/Main.C + examples/Gravity.C
// restricted to the non-periodic GravityVisitor traversal on one GPU.
//
//   pt_gravity -f <tipsy> [-l B] [-e dSoft] [-i iterations] [-v out_prefix]
//              [-m] [-j max_timestep] [-theta t] [-fp32c] [-t oct -d oct]
//
// Per iteration (Driver::run, src/Driver.h:183-288): build the tree, run
// startDown<GravityVisitor>, write the accelerations at iteration 0 when -v is
// given (postIterationFn, examples/Gravity.C:54-61), then kick + perturb with
// timestep min(L / v_max / N^(1/3), max_timestep) (examples/Gravity.C:63-67).
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
    for (size_t i = 0; i < rest.size(); i++) {
      if (rest[i] == "-m") continue;  // Morton keys: the only key of this path
      if (rest[i] == "-j" && i + 1 < rest.size()) max_timestep = std::atof(rest[++i].c_str());
      else if (rest[i] == "-theta" && i + 1 < rest.size()) theta = std::atof(rest[++i].c_str());
      else if (rest[i] == "-fp32c") prec = PT_FP32C;
      else if (rest[i] == "-hexadecapole") conf.hexadecapole = true;  // -DHEXADECAPOLE build
      else {
        std::fprintf(stderr, "unsupported option %s\n", rest[i].c_str());
        return 2;
      }
    }
    if (conf.input_file.empty()) {
      std::fprintf(stderr, "usage: %s -f <tipsy> [-l B] [-e dSoft] [-i it] [-v out] [-theta t] [-fp32c] [-hexadecapole]\n",
                   argv[0]);
      return 2;
    }
    TipsyData t = readTipsy(conf.input_file);
    Driver d(conf);
    d.load(t.pos, t.vel, t.mass, t.soft);
    const int64_t n = d.size();
    const double zero[3] = {0, 0, 0};
    for (int iter = 0; iter < conf.num_iterations; iter++) {
      const pt_counters c = d.computeForces(GravityVisitor(zero, theta), prec);
      std::printf("* Iteration %d: %llu node-particle interactions, %llu bucket-particle interactions "
                  "%llu node opens, %llu node closes\n",
                  iter, (unsigned long long)c.n_node_ints, (unsigned long long)c.n_part_ints,
                  (unsigned long long)c.n_opens, (unsigned long long)c.n_closes);
      if (iter == 0 && !conf.output_file.empty()) d.outputParticleAccelerations(conf.output_file);
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
      d.kickDrift(dt);
    }
    return 0;
  } catch (const paratreet::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
    return 1;
  }
}

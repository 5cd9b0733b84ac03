// dg_cli.cpp -- dgrun: the reference's command line (run_dgcuda and
// get_input, src/2d/conservation/main.cu:481-598, 600-974) over the C ABI of
// libdgamd.so.
//
// Reference options keep their meaning:
//   -n P      polynomial order (0..5)
//   -t STEPS  number of time steps
//   -T TIME   end time (the last step lands on it)
//   -c TOL    run until the step-to-step change is <= TOL
//   -V N      write output/video/U0-<k>.pos every N steps
//   -e        L2 errors ("L2 U<n> error     = ...") and output/U<n>{error,exact}.msh
//   -b        print the run time
//   -v        one line per step
//   -h        usage
// The physics, problem and flux that the reference selects by compiling a
// plugin directory are options here:
//   --physics advection|euler|swe|maxwell-tm|maxwell-te   (default euler)
//   --problem sinsin|linear|vortex|supersonic|const|bump|pulse|constant|rotatinghill
//   --params a,b,...   problem parameters (see include/dg.h)
//   --flux default|llf|upwind|roe   --integrator rk4|rk2|ssp2|ssp3
//   --cfl X  --limiter  --no-positivity  --gamma X  --g X  --a AX,AY  --eps X
//   --output DIR (default output)  --coeffs FILE (modal coefficients: raw
//   little-endian FP64 in [N][n_p][K] order, with FILE.json describing them)
//   --device D  --volume vector|mfma (the stage's volume integral: vector FP64
//   FMAs, or the FP64 matrix cores for Euler p = 4)
// The mesh file is the last argument, as in the reference.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/dg.h"

namespace {

void usage() {
  std::printf("\nUsage: dgrun [OPTIONS] [MESH]\n");
  std::printf(" Options: [-n] Order of polynomial approximation.\n");
  std::printf("          [-t] Number of timesteps.\n");
  std::printf("          [-T] End time.\n");
  std::printf("          [-v] Verbose.\n");
  std::printf("          [-c] Run to convergence of tolerance TOL\n");
  std::printf("          [-V] Print out every N timesteps.\n");
  std::printf("          [-e] Evaluate the error. Requires an exact solution.\n");
  std::printf("          [-b] Benchmark.\n");
  std::printf("          [-h] Display this message.\n");
  std::printf("          [--physics NAME] [--problem NAME] [--params a,b,..] [--flux NAME]\n");
  std::printf("          [--integrator NAME] [--cfl X] [--limiter] [--no-positivity]\n");
  std::printf("          [--gamma X] [--g X] [--a AX,AY] [--eps X] [--output DIR]\n");
  std::printf("          [--coeffs FILE] [--device D] [--volume vector|mfma]\n");
}

int pick(const char *v, const char *const *names, int n) {
  for (int i = 0; i < n; i++)
    if (std::strcmp(v, names[i]) == 0) return i;
  return -1;
}

struct Opts {
  int p = -1, steps = -1, video = 0, verbose = 0, bench = 0, eval_error = 0, convergence = 0;
  double endtime = -1., tol = 0.;
  std::string out = "output", coeffs, mesh;
};

// write_U main.cu:69-116: one gmsh view of per-vertex values
bool write_view(const std::string &path, int n, const std::vector<double> &V, int K,
                const double *u1, const double *u2, const double *u3) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    return false;
  }
  std::fprintf(f, "View \"U%i \" {\n", n);
  for (int e = 0; e < K; e++)
    std::fprintf(f,
                 "ST (%.015lf,%.015lf,0,%.015lf,%.015lf,0,%.015lf,%.015lf,0) "
                 "{%.015lf,%.015lf,%.015lf};\n",
                 V[e], V[(size_t)K + e], V[2 * (size_t)K + e], V[3 * (size_t)K + e],
                 V[4 * (size_t)K + e], V[5 * (size_t)K + e], u1[e], u2[e], u3[e]);
  std::fprintf(f, "};");
  std::fclose(f);
  return true;
}

// the element vertices as read_mesh reads them (main.cu:237-240)
bool read_vertices(const std::string &path, std::vector<double> &V, int &K) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f || std::fscanf(f, "%d", &K) != 1 || K <= 0) {
    if (f) std::fclose(f);
    return false;
  }
  V.assign(6 * (size_t)K, 0.);
  for (int e = 0; e < K; e++) {
    double v[6];
    int s[3];
    if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4],
                    &v[5], &s[0], &s[1], &s[2]) != 9) {
      std::fclose(f);
      return false;
    }
    for (int q = 0; q < 6; q++) V[(size_t)q * K + e] = v[q];
  }
  std::fclose(f);
  return true;
}

int fail(dg_handle *h, const char *what) {
  std::fprintf(stderr, "Error: %s: %s\n", what, h ? dg_last_error(h) : "no handle");
  if (h) dg_destroy(h);
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  dg_config cfg;
  dg_config_default(&cfg);
  Opts o;
  bool have_cfl = false;
  static const char *const phys_names[] = {"advection", "euler", "swe", "maxwell-tm", "maxwell-te"};
  static const char *const prob_names[] = {"sinsin", "linear", "vortex",   "supersonic",  "const",
                                           "bump",   "pulse",  "constant", "rotatinghill"};
  static const char *const flux_names[] = {"default", "llf", "upwind", "roe"};
  static const char *const integ_names[] = {"", "rk4", "rk2", "ssp2", "ssp3"};
  bool have_problem = false;
  if (argc < 2) {
    usage();
    return 1;
  }
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    const bool has = i + 1 < argc;
    auto need = [&]() -> const char * {
      if (!has) {
        usage();
        std::exit(1);
      }
      return argv[++i];
    };
    if (!std::strcmp(a, "-h")) {
      usage();
      return 1;
    } else if (!std::strcmp(a, "-n")) {
      o.p = std::atoi(need());
      if (o.p < 0 || o.p > 5) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "-t")) {
      o.steps = std::atoi(need());
      if (o.steps < 0) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "-T")) {
      o.endtime = std::atof(need());
      if (o.endtime < 0) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "-V")) {
      o.video = std::atoi(need());
    } else if (!std::strcmp(a, "-c")) {
      o.convergence = 1;
      o.tol = std::atof(need());
    } else if (!std::strcmp(a, "-b")) {
      o.bench = 1;
    } else if (!std::strcmp(a, "-e")) {
      o.eval_error = 1;
    } else if (!std::strcmp(a, "-v")) {
      o.verbose = 1;
    } else if (!std::strcmp(a, "--physics")) {
      if ((cfg.physics = pick(need(), phys_names, 5)) < 0) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "--problem")) {
      if ((cfg.problem = pick(need(), prob_names, 9)) < 0) {
        usage();
        return 1;
      }
      have_problem = true;
    } else if (!std::strcmp(a, "--params")) {
      std::string v = need();
      int q = 0;
      size_t pos = 0;
      while (q < 8 && pos <= v.size()) {
        size_t c = v.find(',', pos);
        cfg.problem_params[q++] = std::atof(v.substr(pos, c - pos).c_str());
        if (c == std::string::npos) break;
        pos = c + 1;
      }
    } else if (!std::strcmp(a, "--flux")) {
      if ((cfg.flux = pick(need(), flux_names, 4)) < 0) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "--integrator")) {
      if ((cfg.integrator = pick(need(), integ_names, 5)) <= 0) {
        usage();
        return 1;
      }
    } else if (!std::strcmp(a, "--cfl")) {
      cfg.cfl = std::atof(need());
      have_cfl = true;
    } else if (!std::strcmp(a, "--limiter")) {
      cfg.limiter = 1;
    } else if (!std::strcmp(a, "--no-positivity")) {
      cfg.positivity = 0;
    } else if (!std::strcmp(a, "--gamma")) {
      cfg.gamma = std::atof(need());
    } else if (!std::strcmp(a, "--g")) {
      cfg.g = std::atof(need());
    } else if (!std::strcmp(a, "--eps")) {
      cfg.eps = std::atof(need());
    } else if (!std::strcmp(a, "--a")) {
      std::string v = need();
      size_t c = v.find(',');
      cfg.ax = std::atof(v.substr(0, c).c_str());
      cfg.ay = c == std::string::npos ? cfg.ax : std::atof(v.substr(c + 1).c_str());
    } else if (!std::strcmp(a, "--output")) {
      o.out = need();
    } else if (!std::strcmp(a, "--coeffs")) {
      o.coeffs = need();
    } else if (!std::strcmp(a, "--device")) {
      cfg.device = std::atoi(need());
    } else if (!std::strcmp(a, "--volume")) {
      const std::string v = need();
      if (v == "vector") {
        cfg.volume = 1;
      } else if (v == "mfma") {
        cfg.volume = 2;
      } else {
        usage();
        return 1;
      }
    } else if (i == argc - 1 && a[0] != '-') {
      o.mesh = a;
    } else {
      usage();
      return 1;
    }
  }
  if (o.p < 0 || o.mesh.empty() || (o.steps < 0 && o.endtime < 0 && !o.convergence)) {
    usage();
    return 1;
  }
  cfg.p = o.p;
  if (!have_problem) {
    static const int dflt[] = {DG_PROB_ADV_SINSIN, DG_PROB_EULER_VORTEX, DG_PROB_SWE_BUMP,
                               DG_PROB_MAXWELL_PULSE, DG_PROB_MAXWELL_PULSE};
    cfg.problem = dflt[cfg.physics];
  }
  if (!have_cfl) cfg.cfl = 0.;
  // the reference's sanity check on the limiter (main.cu:647-650)
  if (cfg.limiter && o.p != 1) {
    std::printf("Error: limiter only enabled for p = 1\n");
    return 0;
  }
  std::vector<double> V;
  int K = 0;
  if (!read_vertices(o.mesh, V, K)) {
    std::printf("\nERROR: mesh file not found.\n");
    return 1;
  }
  dg_handle *h = nullptr;
  const int rc0 = dg_create(&h, &cfg);
  if (rc0 != DG_OK) {
    std::fprintf(stderr, "Error: dg_create: %s (code %d)\n",
                 rc0 == DG_ERR_HIP ? "no usable HIP device" : "invalid configuration", rc0);
    return 1;
  }
  if (dg_load_pmsh(h, o.mesh.c_str()) != DG_OK) return fail(h, "read_mesh");
  if (dg_project_ic(h) != DG_OK) return fail(h, "init_conditions");
  dg_info_t info;
  dg_info(h, &info);
  std::printf(" ? %i degree polynomial interpolation (local_n_p = %i)\n", o.p, info.n_p);
  std::printf(" ? %i precomputed basis points\n", info.n_q * info.n_p);
  std::printf(" ? %i elements\n", info.K);
  std::printf(" ? %i sides\n", info.S);
  std::printf(" ? min radius = %.015lf\n", info.min_r);
  if (o.endtime == -1 && !o.convergence)
    std::printf(" ? total_timesteps = %i\n", o.steps);
  else if (o.endtime != -1 && !o.convergence)
    std::printf(" ? endtime = %lf\n", o.endtime);
  else
    std::printf(" ? convergence = %lf\n", o.tol);

  mkdir(o.out.c_str(), 0755);
  const std::string vdir = o.out + "/video";
  if (o.video > 0) mkdir(vdir.c_str(), 0755);
  std::vector<double> u1(K), u2(K), u3(K);
  auto write_u = [&](const std::string &path) -> bool {
    if (dg_eval_u(h, 0, u1.data(), u2.data(), u3.data()) != DG_OK) return false;
    return write_view(path, 0, V, K, u1.data(), u2.data(), u3.data());
  };

  std::printf("Computing...\n");
  auto t0 = std::chrono::steady_clock::now();
  double t = 0.;
  int done = 0, vidnum = 0;
  double conv = 1.;
  const double t_end = o.endtime >= 0 && !o.convergence ? o.endtime : HUGE_VAL;
  const double tol = o.convergence ? o.tol : -1.;
  if (o.video > 0) write_u(vdir + "/U0-" + std::to_string(vidnum++) + ".pos");
  // the loop runs on the device; the host only returns between chunks (one
  // step with -v, the video interval with -V, else the whole run)
  for (;;) {
    int chunk = o.steps >= 0 ? o.steps - done : -1;
    if (o.verbose) chunk = 1;
    if (o.video > 0) chunk = (chunk < 0 || chunk > o.video) ? o.video : chunk;
    if (o.steps >= 0 && done >= o.steps) break;
    int n = 0;
    const double tprev = t;
    int rc = dg_run(h, chunk, t_end, tol, 0., &t, &n, &conv);
    if (rc == DG_ERR_NAN) {
      std::printf("Error: dt is NaN. Dumping...\n");
      break;
    }
    if (rc != DG_OK) return fail(h, "time integration");
    done += n;
    if (o.verbose && n > 0)
      std::printf("(%i) t = %lf, dt = %lf\n", done, t, t - tprev);
    else if (o.convergence)
      std::printf("\r(%i) t = %lf, convergence = %.015lf", done, t, conv);
    else
      std::printf("\r(%i) t = %lf", done, t);
    std::fflush(stdout);
    if (o.video > 0 && n > 0 && done % o.video == 0)
      write_u(vdir + "/U0-" + std::to_string(vidnum++) + ".pos");
    if (n == 0 || n < chunk || chunk < 0) break;  // a stop test ended the run
    if (o.convergence && !(conv > o.tol)) break;
  }
  std::printf("\n");
  if (o.bench) {
    dg_sync(h);
    const double el =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Runtime: %lf seconds\n", el);
  }
  const int N = info.N;
  if (o.eval_error) {
    std::vector<double> err(N);
    if (dg_l2_error(h, t, err.data()) != DG_OK) return fail(h, "eval_error_L2");
    for (int n = 0; n < N; n++) std::printf("L2 U%i error     = %e\n", n, err[n]);
  }
  if (!write_u(o.out + "/U0-final.pos")) return fail(h, "write_U");
  if (o.eval_error) {
    std::vector<double> e1(K), e2(K), e3(K);
    for (int n = 0; n < N; n++) {  // plot_error: u_h - u_exact at the vertices
      if (dg_eval_u(h, n, u1.data(), u2.data(), u3.data()) != DG_OK ||
          dg_eval_exact(h, n, t, e1.data(), e2.data(), e3.data()) != DG_OK)
        return fail(h, "plot_error");
      for (int e = 0; e < K; e++) {
        u1[e] -= e1[e];
        u2[e] -= e2[e];
        u3[e] -= e3[e];
      }
      const std::string p = o.out + "/U" + std::to_string(n) + "error.msh";
      std::printf("Writing to %s...\n", p.c_str());
      write_view(p, n, V, K, u1.data(), u2.data(), u3.data());
    }
    for (int n = 0; n < N; n++) {  // plot_exact
      if (dg_eval_exact(h, n, t, e1.data(), e2.data(), e3.data()) != DG_OK)
        return fail(h, "plot_exact");
      const std::string p = o.out + "/U" + std::to_string(n) + "exact.msh";
      std::printf("Writing to %s...\n", p.c_str());
      write_view(p, n, V, K, e1.data(), e2.data(), e3.data());
    }
  }
  if (!o.coeffs.empty()) {
    std::vector<double> C((size_t)info.dofs);
    if (dg_get_coeffs(h, C.data(), DG_HOST) != DG_OK) return fail(h, "dg_get_coeffs");
    FILE *f = std::fopen(o.coeffs.c_str(), "wb");
    if (!f || std::fwrite(C.data(), sizeof(double), C.size(), f) != C.size()) {
      std::fprintf(stderr, "cannot write %s\n", o.coeffs.c_str());
      if (f) std::fclose(f);
      dg_destroy(h);
      return 1;
    }
    std::fclose(f);
    FILE *j = std::fopen((o.coeffs + ".json").c_str(), "w");
    if (j) {
      std::fprintf(j,
                   "{\"K\": %d, \"p\": %d, \"N\": %d, \"n_p\": %d, \"layout\": "
                   "\"C[n*K*n_p + i*K + e], float64 little-endian\", \"t\": %.17g, "
                   "\"steps\": %d}\n",
                   K, o.p, N, info.n_p, t, done);
      std::fclose(j);
    }
  }
  dg_destroy(h);
  return 0;
}

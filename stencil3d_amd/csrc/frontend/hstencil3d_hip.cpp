// hstencil3d_hip -- the benchmark driver of stencil3d on the GPU.
//
// Keeps the reference driver's command line, phases and report
// (main.cpp:94-332, aux.cpp:8-81) and calls libstencil_hip through its C ABI
// in place of ITERATE / init_arrays / check:
//   init_arrays(ov,os,ou,..)     main.cpp:289   -> sth_init_const(v=0, s=1, u=1)
//   ITERATE(..., 1U, ...)        main.cpp:297   -> sth_iterate(ctx, 1)   (warm-up)
//   check(v,grid,1U)             main.cpp:302   -> sth_check(ctx, 1)
//   ITERATE(..., maxiter, ...)   main.cpp:307   -> sth_iterate(ctx, maxiter), timed
//   check(v,grid,maxiter+1)      main.cpp:312   -> sth_check(ctx, maxiter+1)
// Added options select what the reference fixes at compile time
// (--stencil = STENCIL_SIZE, config.h:37-39; --dtype fp64 = -DUSE_DOUBLE,
// config.h:96-100) and the forms the reference does not have (--mode
// leapfrog, --dtype fp16/bf16, --tune).  --host-fields keeps the reference's
// host arrays (main.cpp:244-289: the --xh/--yh/--zh halos, init_arrays over
// the whole allocation, check on the host) and moves them through
// sth_upload / sth_download.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <vector>

#include "stencil_hip.h"

namespace {

double seconds() {  // aux.cpp:8-13
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + 1e-6 * tv.tv_usec;
}

// aux.cpp:15-43: value range and modulo checks, exit(2) on violation.
int get_int_arg(char *argv[], int iarg, int argc, const char *opt, unsigned &val,
                unsigned minval, unsigned maxval, unsigned modval) {
  if (strcmp(argv[iarg], opt) != 0) return 0;
  if (iarg + 1 >= argc) {
    fprintf(stderr, "%s requires a numeric argument [%u - %u]\n", opt, minval, maxval);
    exit(2);
  }
  val = (unsigned)atoi(argv[iarg + 1]);
  if (val < minval) {
    fprintf(stderr, "%s must be larger or equal to %u\n", opt, minval);
    exit(2);
  }
  if (val > maxval) {
    fprintf(stderr, "%s must be smaller or equal to %u\n", opt, maxval);
    exit(2);
  }
  if (val % modval != 0) {
    fprintf(stderr, "%s must be a multiple of %u\n", opt, modval);
    exit(2);
  }
  printf("%s set to %u\n", opt, val);
  return 2;
}

// aux.cpp:45-69
int get_real_arg(char *argv[], int iarg, int argc, const char *opt, double &val, double minval,
                 double maxval) {
  if (strcmp(argv[iarg], opt) != 0) return 0;
  if (iarg + 1 >= argc) {
    fprintf(stderr, "%s requires a float argument [%f - %f]\n", opt, minval, maxval);
    exit(2);
  }
  val = atof(argv[iarg + 1]);
  if (val < minval) {
    fprintf(stderr, "%s must be larger or equal to %f\n", opt, minval);
    exit(2);
  }
  if (val > maxval) {
    fprintf(stderr, "%s must be smaller or equal to %f\n", opt, maxval);
    exit(2);
  }
  printf("%s set to %f\n", opt, val);
  return 2;
}

// aux.cpp:72-81
int get_flag(char *argv[], int iarg, const char *opt, unsigned &flag) {
  if (strcmp(argv[iarg], opt) != 0) return 0;
  flag = 1;
  printf("%s set to ENABLED\n", opt);
  return 1;
}

int get_choice(char *argv[], int iarg, int argc, const char *opt, const char *const *names,
               int n, int &val) {
  if (strcmp(argv[iarg], opt) != 0) return 0;
  if (iarg + 1 < argc)
    for (int q = 0; q < n; q++)
      if (strcmp(argv[iarg + 1], names[q]) == 0) {
        val = q;
        printf("%s set to %s\n", opt, names[q]);
        return 2;
      }
  fprintf(stderr, "%s requires one of:", opt);
  for (int q = 0; q < n; q++) fprintf(stderr, " %s", names[q]);
  fprintf(stderr, "\n");
  exit(2);
}

void die(const char *what, int st) {
  fprintf(stderr, "ERROR: %s: %s\n", what, sth_strerror(st));
  exit(st == STH_ERR_ARG ? 2 : 3);
}

// The reference's three host arrays (main.cpp:244-275): one dense block per
// field in the s_grid layout with the command line's halos, 64 KiB aligned.
template <class T>
struct HostFields {
  sth_host_layout hl{};
  uint64_t n = 0;
  int64_t jm = 0, im = 0, ofs = 0;
  T *f[3] = {nullptr, nullptr, nullptr};  // indexed by sth_field: v, s, u
  bool alloc(const sth_geom &g, unsigned xh, unsigned yh, unsigned zh) {
    hl = {xh, yh, zh};
    if (sth_host_elems(&g, &hl, &n)) return false;
    jm = (int64_t)g.zs + 2 * zh;
    im = jm * ((int64_t)g.ys + 2 * yh);
    ofs = (int64_t)zh + (int64_t)yh * jm + (int64_t)xh * im;
    const size_t bytes = ((n * sizeof(T) + 65535) / 65536) * 65536;
    for (int q = 0; q < 3; q++)
      if (!(f[q] = (T *)aligned_alloc(65536, bytes))) return false;
    return true;
  }
  ~HostFields() {
    for (int q = 0; q < 3; q++) free(f[q]);
  }
  // init_arrays (iterators.cpp:292-301): every element of the allocation
  void init_arrays() {
    for (uint64_t t = 0; t < n; t++) {
      f[STH_FIELD_V][t] = (T)0;
      f[STH_FIELD_S][t] = (T)1;
      f[STH_FIELD_U][t] = (T)1;
    }
  }
  // check (iterators.cpp:304-325): interior points with |v - expected| > 0.1
  uint64_t check(const sth_geom &g, double expected) const {
    uint64_t err = 0;
    for (int64_t i = 0; i < (int64_t)g.xs; i++)
      for (int64_t j = 0; j < (int64_t)g.ys; j++) {
        const T *row = f[STH_FIELD_V] + ofs + i * im + j * jm;
        for (int64_t k = 0; k < (int64_t)g.zs; k++)
          if (fabs((double)row[k] - expected) > 0.1) err++;
      }
    return err;
  }
  double maxabs(const sth_geom &g, int field) const {
    double m = 0.0;
    for (int64_t i = 0; i < (int64_t)g.xs; i++)
      for (int64_t j = 0; j < (int64_t)g.ys; j++) {
        const T *row = f[field] + ofs + i * im + j * jm;
        for (int64_t k = 0; k < (int64_t)g.zs; k++) m = fmax(m, fabs((double)row[k]));
      }
    return m;
  }
};

template <class T>
int upload_host(sth_ctx *ctx, HostFields<T> &h, int field);
template <>
int upload_host<float>(sth_ctx *ctx, HostFields<float> &h, int field) {
  return sth_upload(ctx, (sth_field)field, h.f[field], &h.hl);
}
template <>
int upload_host<double>(sth_ctx *ctx, HostFields<double> &h, int field) {
  return sth_upload_f64(ctx, (sth_field)field, h.f[field], &h.hl);
}
template <class T>
int download_host(sth_ctx *ctx, HostFields<T> &h, int field);
template <>
int download_host<float>(sth_ctx *ctx, HostFields<float> &h, int field) {
  return sth_download(ctx, (sth_field)field, h.f[field], &h.hl);
}
template <>
int download_host<double>(sth_ctx *ctx, HostFields<double> &h, int field) {
  return sth_download_f64(ctx, (sth_field)field, h.f[field], &h.hl);
}

}  // namespace

int main(int argc, char **argv) {
  // defaults of main.cpp:96-113 / config.h:4-17
  unsigned train = 0, offset = 16, maxiter = 100, cacheiter = 1, dups = 1, prefetch = 64;
  unsigned xs = 945, ys = 945, zs = 880, xh = 4, yh = 4, zh = 16;
  unsigned xb = 1, yb = 1, zb = 16;
  double sbm = 12.0;
  unsigned npts = 7, seed = 0, tune = 0, rows = 0, waves = 0, chunks = 0, vec = 0;
  unsigned spp = 0, launch = 0, rsweeps = 0, host_fields = 0;
  int mode = 0, dtype = 0, init = -1;
  double c1 = 0.25, c2 = 1e-6, c0 = -1e30;
  static const char *const kModes[] = {"ref", "leapfrog"};
  static const char *const kTypes[] = {"fp32", "fp16", "bf16", "fp64"};
  static const char *const kInits[] = {"const", "uniform", "gauss"};

  int iarg = 1;
  while (iarg < argc) {
    int args = 0;
    args += get_int_arg(argv, iarg, argc, "--iter", maxiter, 1, 100000, 1);
    args += get_int_arg(argv, iarg, argc, "--citer", cacheiter, 1, 100000, 1);
    args += get_int_arg(argv, iarg, argc, "--dups", dups, 1, 64, 1);
    args += get_int_arg(argv, iarg, argc, "--ofs", offset, 0, 16384, 1);
    args += get_int_arg(argv, iarg, argc, "--pref", prefetch, 1, 512, 1);
    args += get_int_arg(argv, iarg, argc, "--xh", xh, 4, 16, 1);
    args += get_int_arg(argv, iarg, argc, "--yh", yh, 4, 16, 1);
    args += get_int_arg(argv, iarg, argc, "--zh", zh, 1, 16, 1);
    args += get_int_arg(argv, iarg, argc, "--xs", xs, 1, 1264, 1);
    args += get_int_arg(argv, iarg, argc, "--ys", ys, 1, 1264, 1);
    args += get_int_arg(argv, iarg, argc, "--zs", zs, 1, 1264, 1);
    args += get_int_arg(argv, iarg, argc, "--xb", xb, 1, 1264, 1);
    args += get_int_arg(argv, iarg, argc, "--yb", yb, 1, 1264, 1);
    args += get_int_arg(argv, iarg, argc, "--zb", zb, 1, 1264, 1);
    args += get_real_arg(argv, iarg, argc, "--sbm", sbm, 1.0, 10.0);
    args += get_flag(argv, iarg, "--train", train);
    // GPU build options
    args += get_int_arg(argv, iarg, argc, "--stencil", npts, 7, 25, 1);
    args += get_choice(argv, iarg, argc, "--mode", kModes, 2, mode);
    args += get_choice(argv, iarg, argc, "--dtype", kTypes, 4, dtype);
    args += get_choice(argv, iarg, argc, "--init", kInits, 3, init);
    args += get_int_arg(argv, iarg, argc, "--seed", seed, 0, 1u << 30, 1);
    args += get_real_arg(argv, iarg, argc, "--c1", c1, 0.0, 4.0);
    args += get_real_arg(argv, iarg, argc, "--c2", c2, -1e6, 1e6);
    args += get_real_arg(argv, iarg, argc, "--c0", c0, -1e6, 1e6);
    args += get_flag(argv, iarg, "--tune", tune);
    args += get_int_arg(argv, iarg, argc, "--rows", rows, 1, 8, 1);
    args += get_int_arg(argv, iarg, argc, "--waves", waves, 1, 16, 1);
    args += get_int_arg(argv, iarg, argc, "--chunks", chunks, 1, 64, 1);
    args += get_int_arg(argv, iarg, argc, "--vec", vec, 1, 2, 1);
    args += get_int_arg(argv, iarg, argc, "--steps-per-pass", spp, 1, 5, 1);
    args += get_int_arg(argv, iarg, argc, "--launch", launch, 1, 2, 1);
    args += get_int_arg(argv, iarg, argc, "--sweeps-per-launch", rsweeps, 1, 4096, 1);
    args += get_flag(argv, iarg, "--host-fields", host_fields);
    if (args > 0) {
      iarg += args;
    } else {
      fprintf(stderr, "unsupported argument %s\n", argv[iarg]);
      return 2;
    }
  }
  if (npts != 7 && npts != 13 && npts != 25) {
    fprintf(stderr, "--stencil must be 7, 13 or 25\n");
    return 2;
  }
  if ((ys % yb) != 0) {  // main.cpp:171-174
    fprintf(stderr, "ERROR: --ys must by an exact multiple of --yb\n");
    exit(2);
  }
  if (mode == 0 && dtype != 0 && dtype != 3) {
    fprintf(stderr, "ERROR: --mode ref runs FP32 or FP64 storage only\n");
    exit(2);
  }
  if (host_fields && dtype != 0 && dtype != 3) {
    fprintf(stderr, "ERROR: --host-fields needs FP32 or FP64 storage\n");
    exit(2);
  }
  if (init < 0) init = mode == 0 ? 0 : 1;
  const unsigned R = npts == 7 ? 1 : npts == 13 ? 2 : 4;
  if (xh < R || yh < R || zh < R) {
    fprintf(stderr, "ERROR: halos must be at least the stencil radius %u\n", R);
    exit(2);
  }
  // --yb maps onto the tile height (rows per wave x waves) when it is one of
  // the compiled tiles; --zb is fixed at one wave (64 k) per row on the GPU.
  if (!rows && !waves && yb > 1) {
    static const unsigned cand[][2] = {{1, 4}, {2, 4}, {4, 4}, {8, 4}, {4, 8}, {8, 8}};
    for (auto &c : cand)
      if (c[0] * c[1] == yb) {
        rows = c[0];
        waves = c[1];
      }
  }

  // the report's constants (config.h:73-91)
  const double flops_pt = npts == 7 ? 12 : npts == 13 ? 29 : 41;
  const double cache_pt = npts == 7 ? 40 : npts == 13 ? 64 : 112;
  const double eb = dtype == 0 ? 4.0 : dtype == 3 ? 8.0 : 2.0;
  const double mem_pt = 4 * eb;
  const double gvol = (double)xs * ys * zs;
  const double flops = flops_pt * 1.0e-9 * gvol * maxiter * cacheiter;
  const double mem_bytes = mem_pt * 1.0e-9 * gvol * maxiter * cacheiter;
  const double cache_bytes = cache_pt * 1.0e-9 * gvol * maxiter * cacheiter;
  int phase = 0;

  printf("running STENCIL(%u) benchmark: \n", npts);
  sth_geom g = {xs, ys, zs, npts, (uint32_t)dtype};
  sth_layout_info L;
  int st = sth_layout(&g, &L);
  if (st) die("layout", st);
  printf(" -- using a grid size of %ux%ux%u with a halo of %ux%ux%u elements totaling %f GB\n", xs,
         ys, zs, R, R, R, 1e-9 * (double)L.bytes);
  printf(" -- device layout: row pitch %llu, plane %llu elements, %s storage, 256-B aligned rows\n",
         (unsigned long long)L.pitch, (unsigned long long)L.plane, kTypes[dtype]);
  printf(" -- scheduling: 2.5D sweep along x over (y,z) tiles, HIP on gfx950\n");

  sth_ctx *ctx = nullptr;
  if ((st = sth_create(&g, &ctx))) die("sth_create", st);
  sth_tuning tn = {rows, waves, chunks, 1, vec, spp, launch, rsweeps};
  if ((st = sth_set_tuning(ctx, &tn))) die("tuning", st);
  // FP64 is the reference's -DUSE_DOUBLE build: its weights too
  if (dtype == 3 && (st = sth_set_weights(ctx, STH_WEIGHTS_REFERENCE))) die("weights", st);

  HostFields<float> hf32;
  HostFields<double> hf64;
  if (host_fields) {
    const bool ok = dtype == 3 ? hf64.alloc(g, xh, yh, zh) : hf32.alloc(g, xh, yh, zh);
    if (!ok) die("host arrays", STH_ERR_NOMEM);
    printf(" -- host arrays: halo %ux%ux%u, %f GB per field\n", xh, yh, zh,
           1e-9 * (double)(dtype == 3 ? hf64.n * 8 : hf32.n * 4));
  }
  // Device fields from the host arrays (the reference's whole-allocation
  // init_arrays, or a device initialiser round-tripped through the host
  // layout).
  auto via_host = [&](bool const_init) {
    int s2 = 0;
    for (int q = 0; q < 3 && !s2; q++) {
      if (dtype == 3) {
        if (const_init) {
          if (q == 0) hf64.init_arrays();
        } else {
          s2 = download_host(ctx, hf64, q);
        }
        if (!s2) s2 = upload_host(ctx, hf64, q);
      } else {
        if (const_init) {
          if (q == 0) hf32.init_arrays();
        } else {
          s2 = download_host(ctx, hf32, q);
        }
        if (!s2) s2 = upload_host(ctx, hf32, q);
      }
    }
    return s2;
  };
  auto initialise = [&]() {
    int s2;
    if (init == 0) {
      if (host_fields) return via_host(true);
      s2 = sth_init_const(ctx, 0.0f, 1.0f, 1.0f);  // init_arrays, iterators.cpp:292-301
    } else if (init == 1) {
      s2 = sth_init_uniform(ctx, STH_FIELD_U, seed, 0, -1.0, 1.0, 0.0f);
      if (!s2) s2 = sth_init_uniform(ctx, STH_FIELD_S, seed, 1, -1.0, 1.0, 0.0f);
      if (!s2) s2 = sth_init_uniform(ctx, STH_FIELD_V, seed, 2, -1.0, 1.0, 0.0f);
    } else {
      s2 = sth_init_gaussian(ctx, STH_FIELD_U, 1.0, 8.0, 1e-3, seed, 0);
      if (!s2) s2 = sth_init_gaussian(ctx, STH_FIELD_V, 1.0, 8.0, 1e-3, seed, 1);
      if (!s2) s2 = sth_init_point(ctx, STH_FIELD_S, 0, 0, 0, 0.0f);
    }
    if (!s2 && host_fields) s2 = via_host(false);
    if (!s2) s2 = sth_sync(ctx);
    return s2;
  };

  printf("[rank0] Phase #%i: array initialization\n", ++phase);
  double t0 = seconds();
  if ((st = initialise())) die("initialization", st);
  double t1 = seconds();
  printf("[rank0] Phase #%i: initialization took %f secs\n", ++phase, t1 - t0);

  if (tune) {
    sth_tuning best;
    float ms = 0.0f;
    if ((st = sth_autotune(ctx, mode, 3, &best, &ms))) die("autotune", st);
    printf(" -- autotune: rows/wave %u, waves %u, k/lane %u, chunks %u: %.3f ms/sweep\n",
           best.rows_per_wave, best.waves, best.vec, best.chunks, ms);
    // the tuning sweeps modified the fields (and 16-bit scales): the timed
    // run starts from the same state as without --tune
    if ((st = initialise())) die("re-initialization", st);
  }

  const float fc1 = (float)c1;
  // default centre weight: 2 - c1 * (the neighbour sum's constant-mode gain),
  // which keeps a constant field constant (6 for 7 and 13 points, 12 for 25)
  float c0_default = 0.0f;
  if ((st = sth_leapfrog_c0(npts, fc1, &c0_default))) die("c0", st);
  const float fc0 = c0 > -1e29 ? (float)c0 : c0_default;
  std::vector<float> c2v(maxiter + 1, (float)c2);
  auto sweep = [&](unsigned n) {
    return mode == 0 ? sth_iterate(ctx, n) : sth_leapfrog(ctx, n, fc0, fc1, c2v.data());
  };
  auto report_check = [&](unsigned expected) {
    if (mode != 0) return;
    uint64_t errors = 0;
    int s2 = 0;
    if (host_fields) {
      // the reference's check on its host array (iterators.cpp:304-325)
      s2 = dtype == 3 ? download_host(ctx, hf64, STH_FIELD_V) : download_host(ctx, hf32, STH_FIELD_V);
      if (!s2) errors = dtype == 3 ? hf64.check(g, expected) : hf32.check(g, expected);
    } else {
      s2 = sth_check(ctx, (float)expected, &errors);
    }
    if (s2) die("check", s2);
    printf("check: %llu errors found\n", (unsigned long long)errors);
  };

  printf("[rank0] Phase #%i: initial warmup\n", ++phase);
  t0 = seconds();
  if ((st = sweep(1)) || (st = sth_sync(ctx))) die("warmup", st);
  t1 = seconds();
  printf("[rank0] Phase #%i: finish warmup in %f secs\n", ++phase, t1 - t0);
  report_check(1);
  // LEAPFROG: the scheme's discrete energy before the timed steps (FP64 on
  // the GPU, untimed); with s = 0 and zero halos the steps conserve it
  double q0 = 0.0;
  if (mode != 0 && (st = sth_energy(ctx, fc0, fc1, &q0))) die("energy", st);

  printf("[rank0] Phase #%i: start compute\n", ++phase);
  if ((st = sth_timer_start(ctx))) die("timer", st);
  t0 = seconds();
  if ((st = sweep(maxiter)) || (st = sth_sync(ctx))) die("compute", st);
  t1 = seconds();
  double kms = 0.0;
  uint64_t launches = 0;
  if ((st = sth_timer_read(ctx, &kms, &launches))) die("timer", st);
  // main.cpp:309-310
  printf("[rank0] Phase #%i: finish compute in %f secs (%f %s, mem: %f GB/s, cache: %f GB/s)\n",
         ++phase, t1 - t0, flops / (t1 - t0), "GFlop/s", mem_bytes / (t1 - t0) / cacheiter,
         cache_bytes / (t1 - t0));
  // kernel time per sweep: a launch may apply several (fused LEAPFROG
  // passes, REF sweeps per launch), and each launch streams at least the
  // four fields of one sweep (three read, one written) through HBM
  const double per = kms / (double)(maxiter ? maxiter : 1);
  const double gpts = gvol / (per * 1e-3) * 1e-9;
  const double hbm = kms > 0 ? mem_pt * gvol * (double)launches / (kms * 1e-3) * 1e-9 : 0.0;
  printf(" -- GPU: %.4f ms/sweep (HIP events, %llu launches), %.2f Gpts/s, %.1f GB/s "
         "HBM at four fields per launch (%.1f%% of 8 TB/s)\n",
         per, (unsigned long long)launches, gpts, hbm, 100.0 * hbm / 8000.0);
  report_check(maxiter + 1);
  if (mode != 0) {
    double mx = 0, l2 = 0;
    if ((st = sth_norms(ctx, STH_FIELD_U, &mx, &l2))) die("norms", st);
    printf("norms: max |u| = %.9e, ||u||_2 = %.9e\n", mx, l2);
    double q1 = 0.0;
    if ((st = sth_energy(ctx, fc0, fc1, &q1))) die("energy", st);
    printf("energy: before %.9e, after %.9e, relative change %.3e\n", q0, q1,
           q0 != 0.0 ? fabs(q1 - q0) / fabs(q0) : fabs(q1 - q0));
    if (host_fields) {
      const int s2 = dtype == 3 ? download_host(ctx, hf64, STH_FIELD_U)
                                : download_host(ctx, hf32, STH_FIELD_U);
      if (s2) die("download", s2);
      printf("host: max |u| = %.9e\n", dtype == 3 ? hf64.maxabs(g, STH_FIELD_U)
                                                   : hf32.maxabs(g, STH_FIELD_U));
    }
  }
  printf("freeing data arrays\n");
  sth_destroy(ctx);
  printf("[rank0] Phase #%i: done!\n", ++phase);
  (void)train;
  (void)offset;
  (void)dups;
  (void)prefetch;
  (void)xb;
  (void)zb;
  (void)sbm;
  return 0;
}

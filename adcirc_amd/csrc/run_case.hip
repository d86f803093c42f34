// adcgpu_run_case: runs a case written by adcirc_amd.model.write_case through
// the C ABI only (no Python in the process), for profilers and as a worked
// example of the call sequence a host model makes:
//
//   adcgpu_create -> adcgpu_cold_start -> adcgpu_holland (time 0)
//   -> adcgpu_run (warm-up steps) -> adcgpu_run (timed steps)
//
//   adcgpu_run_case CASE [--warmup W] [--steps K] [--chunk C] [--convcr TOL]
//                   [--precision 0|1] [--copy-mib N]
//   adcgpu_run_case --copy-mib N
//
// Prints one line of timing: steps, ms per step, JCG iterations per step.
// --convcr replaces the case's JCG tolerance (a different iteration count per
// launch, for fitting per-iteration and per-launch HBM traffic); --precision
// selects ADCGPU_FP64 (0) or ADCGPU_MIXED (1), --warm-start the JCG start.
// --state FILE --it0 S starts from a snapshot (adcgpu_state_save) taken
// after step S instead of the cold start, and --params FILE loads the
// learned start's ADCGPU_ML_NPARAM float32 parameters (adcgpu_ml_params);
// bench.py --dump-legs writes all three for the legs it times.
// --copy-mib runs a plain device-to-device copy kernel over N MiB ten times
// (after the case, when one is given):
// the reference reading of the HBM traffic counters (FETCH_SIZE, WRITE_SIZE)
// for a kernel whose bytes are known.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/adcgpu.h"

namespace {

__global__ void copy_f64(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int copy_probe(int mib) {
  const size_t n = (size_t)mib * 1024 * 1024 / sizeof(double);
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) return 1;
  if (hipMemset(a, 0, n * 8) != hipSuccess) return 1;
  for (int r = 0; r < 10; ++r) hipLaunchKernelGGL(copy_f64, dim3(4096), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copy kernel: 10 launches of %d MiB read + %d MiB written\n", mib, mib);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

struct Record {
  int kind = 0;  // 0 float64, 1 int32
  std::vector<double> d;
  std::vector<int32_t> i;
};

bool read_case(const char* path, std::map<std::string, Record>& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char magic[8];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "ADCGCASE", 8) != 0) {
    std::fclose(f);
    return false;
  }
  for (;;) {
    char name[17] = {0};
    int32_t kind[2];
    int64_t count = 0;
    if (std::fread(name, 1, 16, f) != 16) break;
    if (std::fread(kind, 4, 2, f) != 2 || std::fread(&count, 8, 1, f) != 1 || count < 0) {
      std::fclose(f);
      return false;
    }
    Record r;
    r.kind = kind[0];
    size_t got = 0;
    if (r.kind == 0) {
      r.d.resize((size_t)count);
      got = std::fread(r.d.data(), 8, (size_t)count, f);
    } else {
      r.i.resize((size_t)count);
      got = std::fread(r.i.data(), 4, (size_t)count, f);
    }
    if (got != (size_t)count) {
      std::fclose(f);
      return false;
    }
    out[name] = std::move(r);
  }
  std::fclose(f);
  return true;
}

const Record* need(const std::map<std::string, Record>& c, const char* k, int kind) {
  auto it = c.find(k);
  if (it == c.end() || it->second.kind != kind) {
    std::fprintf(stderr, "case file: missing or mistyped record %s\n", k);
    std::exit(2);
  }
  return &it->second;
}

int check(int rc, const char* what) {
  if (rc != 0) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, adcgpu_last_error());
    std::exit(1);
  }
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr,
                 "usage: %s CASE [--warmup W] [--steps K] [--chunk C] [--convcr TOL] "
                 "[--precision 0|1] [--warm-start 0|1|2] [--state FILE --it0 S] "
                 "[--params FILE] [--copy-mib N] | --copy-mib N\n",
                 argv[0]);
    return 2;
  }
  if (argc == 3 && !std::strcmp(argv[1], "--copy-mib")) return copy_probe(std::atoi(argv[2]));
  int warmup = 10, steps = 40, chunk = 200, copy_mib = 0, precision = -1, warm = -1, it0 = 0;
  double convcr = 0.0;  // > 0: replaces the case's JCG tolerance
  const char *state_path = nullptr, *params_path = nullptr;
  for (int a = 2; a + 1 < argc; a += 2) {
    if (!std::strcmp(argv[a], "--warmup")) warmup = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--steps")) steps = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--chunk")) chunk = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--copy-mib")) copy_mib = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--convcr")) convcr = std::atof(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--precision")) precision = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--warm-start")) warm = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--state")) state_path = argv[a + 1];
    else if (!std::strcmp(argv[a], "--it0")) it0 = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--params")) params_path = argv[a + 1];
  }
  if (it0 < 0 || (it0 > 0) != (state_path != nullptr)) {
    std::fprintf(stderr, "--state and --it0 go together\n");
    return 2;
  }
  std::map<std::string, Record> c;
  if (!read_case(argv[1], c)) {
    std::fprintf(stderr, "cannot read case file %s\n", argv[1]);
    return 2;
  }
  auto I = [&](const char* k) { return need(c, k, 1)->i[0]; };
  auto D = [&](const char* k) { return need(c, k, 0)->d[0]; };
  auto IA = [&](const char* k) { return need(c, k, 1)->i.data(); };
  auto DA = [&](const char* k) { return need(c, k, 0)->d.data(); };

  adcgpu_mesh m;
  m.np = I("np"); m.ne = I("ne"); m.mnei = I("mnei"); m.neta = I("neta");
  m.nvel = I("nvel"); m.nvelme = I("nvelme");
  m.nm = IA("nm"); m.x = DA("x"); m.y = DA("y"); m.dp = DA("dp");
  m.fdxe = DA("fdxe"); m.fdye = DA("fdye"); m.areas = DA("areas");
  m.nneigh = IA("nneigh"); m.neitab = IA("neitab"); m.csr_position = IA("csr_position");
  m.nbd = IA("nbd"); m.nbv = IA("nbv"); m.lbcodei = IA("lbcodei");
  m.csii = DA("csii"); m.siii = DA("siii"); m.me2gw = IA("me2gw");
  adcgpu_params p;
  p.dtdp = D("dtdp"); p.g = D("g"); p.h0 = D("h0"); p.velmin = D("velmin");
  p.a00 = D("a00"); p.b00 = D("b00"); p.c00 = D("c00"); p.convcr = D("convcr");
  p.prbckgrnd_mh2o = D("prbckgrnd_mh2o");
  p.nolifa = I("nolifa"); p.nolica = I("nolica"); p.nolicat = I("nolicat");
  p.nolibf = I("nolibf"); p.itmax = I("itmax"); p.nws = I("nws"); p.precision = I("precision");
  p.ilump = c.count("ilump") ? I("ilump") : 0;
  p.warm_start = c.count("warm_start") ? I("warm_start") : 0;
  p.errorelev = c.count("errorelev") ? D("errorelev") : 0.0;
  p.tau0var = DA("tau0var"); p.evm = DA("evm"); p.fric = DA("fric"); p.corif = DA("corif");
  if (convcr > 0.0) p.convcr = convcr;
  if (precision >= 0) p.precision = precision;
  if (warm >= 0) p.warm_start = warm;

  const std::vector<double>& table = need(c, "eta_bc_table", 0)->d;
  const int neta = m.neta > 0 ? m.neta : 1;
  if ((int64_t)table.size() < (int64_t)(it0 + warmup + steps) * neta) {
    std::fprintf(stderr, "eta_bc_table holds %zu values, %d steps need %lld\n", table.size(),
                 it0 + warmup + steps, (long long)(it0 + warmup + steps) * neta);
    return 2;
  }
  const double* hol = c.count("holland") ? DA("holland") : nullptr;

  adcgpu_ctx* ctx = nullptr;
  check(adcgpu_create(&m, &p, 0, &ctx), "adcgpu_create");
  check(adcgpu_cold_start(ctx, DA("eta2_0"), IA("nodecode_0"), IA("noff_0")), "cold_start");
  if (hol) check(adcgpu_holland(ctx, 0.0, hol), "holland");
  if (state_path) {
    // a snapshot (adcgpu_state_save) of the same mesh taken after step it0
    std::vector<char> buf((size_t)adcgpu_state_bytes(ctx));
    FILE* f = std::fopen(state_path, "rb");
    if (!f || std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
      std::fprintf(stderr, "cannot read %zu snapshot bytes from %s\n", buf.size(), state_path);
      return 2;
    }
    std::fclose(f);
    check(adcgpu_state_load(ctx, buf.data(), (int64_t)buf.size()), "state_load");
  }
  if (params_path) {
    std::vector<float> prm(ADCGPU_ML_NPARAM);
    FILE* f = std::fopen(params_path, "rb");
    if (!f || std::fread(prm.data(), 4, prm.size(), f) != prm.size()) {
      std::fprintf(stderr, "cannot read %d parameters from %s\n", ADCGPU_ML_NPARAM, params_path);
      return 2;
    }
    std::fclose(f);
    check(adcgpu_ml_params(ctx, prm.data(), nullptr), "ml_params");
  }
  std::fprintf(stderr, "[run_case] %d nodes, %d elements: warm-up %d steps from step %d\n", m.np,
               m.ne, warmup, it0 + 1);
  if (warmup > 0)
    check(adcgpu_run(ctx, it0 + 1, warmup, table.data() + (size_t)it0 * neta, hol, 0), "run");
  check(adcgpu_synchronize(ctx), "synchronize");
  std::fprintf(stderr, "[run_case] timed %d steps\n", steps);
  int64_t iters = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int done = 0; done < steps;) {
    const int k = std::min(chunk, steps - done);
    const int s0 = it0 + warmup + done;
    check(adcgpu_run(ctx, s0 + 1, k, table.data() + (size_t)s0 * neta, hol, 0), "run");
    iters += adcgpu_run_iterations(ctx);
    done += k;
    std::fprintf(stderr, "[run_case]   %d/%d steps\n", done, steps);
  }
  check(adcgpu_synchronize(ctx), "synchronize");
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("steps %d ms_per_step %.4f jcg_iterations_per_step %.3f\n", steps, ms / steps,
              steps ? (double)iters / steps : 0.0);
  std::fflush(stdout);
  adcgpu_destroy(ctx);
  return copy_mib > 0 ? copy_probe(copy_mib) : 0;
}

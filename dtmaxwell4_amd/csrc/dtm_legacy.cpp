// dtm_legacy.cpp — the DTmxw surface that the reference's SWIG module binds
// (mxw.i:1-36) and run.py drives (run.py:11-40): the globals Tsteps and
// shotpoint, the grid constants, TFSFsrc::set/check, _main and calcStep.
//
// Differences from the reference, all deliberate:
//  * grid sizes and steps are run-time globals (no recompile per size) and
//    GridNx/GridNy/GridNz are the true x/y/z extents (the reference's
//    Python-visible GridNy/GridNz are swapped, py_consts.h:10-12 vs
//    params.h:44-49);
//  * the absorbing layer is the reference's split-field layer with its sigma
//    profile (DTM_PML_PROFILE_REFERENCE: the x grading of DTmxw.cu:219-253
//    on every face), Npml* cells per side, on every face but the low-z one
//    (non-absorbing there, as in the reference, DTmxw.cu:262; the low-z face
//    is a zero boundary inside the silver film); the reference's y and z
//    tables call the same sigma function with other index offsets
//    (DTmxw.cu:257, 261), which is not reproduced;
//  * no OpenGL viewer: without --test, _main reports that and returns -1.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/dtmaxwell4.h"

extern "C" {
int Tsteps = 22 * 10;  // DTmxw.cu:730 (Ntime*10)
TFSFsrc shotpoint;
int DropData = 0;  // py_consts.h:19: DROP_DATA is off in the reference's default build
int GridNx = 600, GridNy = 639, GridNz = 300;  // true x, y, z extents of the reference default
float ds = 0.005f, da = 0.005f, dv = 0.005f, dt = 0.002f;  // py_consts.h:71
const int NDT = 3;
int Npmlx = 15, Npmly = 15, Npmlz = 16;  // Npmlx/2 rags x 3 cells, Npmlz/2 cells
int Ntime = 22;                          // NTIME = NS-3 (params.h:50-52)
int NyBloch = 1, Nzw = 300, Np = 200;
}

namespace {

// Materials of init_material (DTmxw.cu:604-642), enum order of defs.h:121.
enum { IndAir = 0, IndBIG, IndGold, IndGGG, IndOut, IndSh1, IndSh2, IndSh3, IndSh4, IndSh5, IndVac, IndAg, IndOrg,
       IndITO, IndGlass, Nmats };

struct LegacyRun {
  dtm_solver* s = nullptr;
  std::string drop_dir = ".";
  struct SensorFile {
    std::string fld;
    int x, y, z;
    FILE* f;
    int probe;
  };
  std::vector<SensorFile> sensors;
  std::vector<int> extra;  // --sensor points (x y z triples)
  int iStep = 0;
};
LegacyRun g_run;

void set_materials(dtm_config& c) {
  c.nmat = Nmats;
  const double n2 = sqrt(5.0), n3 = 2.0;
  const double eps[Nmats] = {1.0, n2 * n2, n2 * n2, n3 * n3, 100.0 * 100.0, 1, 4, 9, 16, 25, 1, 1, 1.75 * 1.75,
                             1.8 * 1.8, 1.5 * 1.5};
  for (int m = 0; m < Nmats; m++) {
    memset(&c.mat[m], 0, sizeof(dtm_material));
    c.mat[m].eps_inf = eps[m];
  }
  // IndGold: two poles (DTmxw.cu:625-630): setDrude(8.5, 4.9pi, 0.21, 1e-3), set(1, pi, 60pi, 340pi)
  c.mat[IndGold].eps_inf = 1.0;
  c.mat[IndGold].npoles = 2;
  LorenzDisp_setDrude(&c.mat[IndGold].pole[0], 8.5, 4.9 * M_PI, 0.21, 1e-3);
  LorenzDisp_set(&c.mat[IndGold].pole[1], 1.0, 0.5 * 2 * M_PI, 30 * 2 * M_PI, 340 * M_PI);
  // IndAg: Drude (DTmxw.cu:632-635)
  c.mat[IndAg].eps_inf = 1.0;
  c.mat[IndAg].npoles = 1;
  LorenzDisp_setDrude(&c.mat[IndAg].pole[0], 1.0, 136.66 / 3.0, 2.69e13 / 2 / 3e14, 0);
}

// get_mat_indev (defs.h:128-148) on half-coordinates (ix, iy, iz).
int mat_at(int ix, int iy, int iz) {
  const float dx = ds, dy = da, dz = dv;
  float x = ix * 0.5f * dx, y = iy * 0.5f * dy, z = iz * 0.5f * dz;
  const float Xc = 0.5f * GridNx * dx, Yc = 0.5f * GridNy * dy;
  const float mx = (2 * Npmlx + 100) * dx / 2, my = (2 * Npmly + 100) * dy / 2;
  if (x < mx || x > GridNx * dx - mx || y < my || y > GridNy * dy - my) return IndOrg;
  const float period = 0.4f, rad = 0.1f, height = 0.1f;
  const float xr = x - Xc - roundf((x - Xc) / period) * period;
  const float yr = y - Yc - roundf((y - Yc) / period) * period;
  const float shift = 1e-5f;
  z -= 0.150f - shift;
  if (z < 0) {
    if (z > -height && xr * xr + yr * yr < rad * rad) return IndOrg;
    return IndAg;
  }
  z -= 0.180f + shift;
  if (z < 0) return IndOrg;
  z -= 0.100f + shift;
  if (z < 0) return IndITO;
  return IndGlass;
}

void build_geometry(std::vector<uint8_t> m[3]) {
  const size_t n = (size_t)GridNx * GridNy * GridNz;
  static const int off[3][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
  for (int c = 0; c < 3; c++) m[c].assign(n, 0);
  for (int i = 0; i < GridNx; i++)
    for (int j = 0; j < GridNy; j++)
      for (int k = 0; k < GridNz; k++) {
        const size_t q = ((size_t)i * GridNy + j) * GridNz + k;
        for (int c = 0; c < 3; c++)
          m[c][q] = (uint8_t)mat_at(2 * i + off[c][0], 2 * j + off[c][1], 2 * k + off[c][2]);
      }
}

int print_help() {
  printf("using: ./DTmxw [--help] [--test] [--drop_dir \".\"] [--sensor \"1 1 1\"] [--swap_dir \".\"]\n");
  printf("  --test\theadless run of Tsteps steps (Tsteps/Ntime calls of calcStep);\n");
  printf("  --drop_dir\tdirectory for sensor files, [.];\n");
  printf("  --sensor\tsensor node \"x y z\": Ex..Hz files <fld>-xXXXXXyYYYYYzZZZZZ.dat;\n");
  printf("  GL options (--box --mesh --Dmesh --zoom --step --bkgr_col --mesh_col --box_col) are accepted and ignored.\n");
  return 0;
}

void add_sensor_files(int x, int y, int z) {
  static const char* f[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int q = 0; q < 6; q++) {
    char fname[512];
    snprintf(fname, sizeof(fname), "%s/%s-x%05dy%05dz%05d.dat", g_run.drop_dir.c_str(), f[q], x, y, z);  // sensor.h:17
    printf("Set sensor %s at %d %d %d\n", f[q], x, y, z);
    FILE* fp = fopen(fname, "w");
    if (!fp) {
      printf("Error openning file %s\n", fname);
      continue;
    }
    const int p = dtm_add_sensor(g_run.s, f[q], x, y, z);
    if (p < 0) {
      fclose(fp);
      continue;
    }
    g_run.sensors.push_back({f[q], x, y, z, fp, p});
  }
}

}  // namespace

extern "C" {

// TFSFsrc::set (signal.hpp:99-115)
void TFSFsrc_set(TFSFsrc* s, double LightSpeed, double impedance, double lmb) {
  const double c = LightSpeed;
  s->Imp = (float)impedance;
  s->V_max = (float)c;
  s->cSrc = (float)c;
  s->wavelength = (float)lmb;
  s->k = (float)(2 * M_PI / s->wavelength);
  s->w = (float)(c * s->k);
  s->Omega = (float)(0.05 * s->w);
  s->kEnv = s->Omega / s->w;
  s->theta = (float)(0 + M_PI);
  s->phi = (float)(M_PI / 2);
  s->t0 = (float)(25 * 0.02);
  s->tw = (float)(5 * 0.02);
  s->T = (float)(125 * 0.02);
}

// TFSFsrc::check (signal.hpp:116-133)
void TFSFsrc_check(TFSFsrc* s) {
  printf("TF/SF source: Box(s,v,a) |%g-%g|x|%g-%g|x|%g-%g|\n", s->BoxMs, s->BoxPs, s->BoxMv, s->BoxPv, s->BoxMa,
         s->BoxPa);
  printf("TF/SF source: Shotpoint(s,v,a) = %g,%g,%g\n", s->srcXs, s->srcXv, s->srcXv);
  printf("TF/SF source: wavelength = %g, kEnv=%g, tStop=%g\n", s->wavelength, s->kEnv, s->tStop);
  printf("TF/SF stops after %d steps\n", (int)(s->tStop / dt));
  if (s->BoxMs < 0) printf("BoxMs<0: incorrect results are possible\n");
  if (s->BoxMa < 0) printf("BoxMa<0: incorrect results are possible\n");
  if (s->BoxMv < 0) printf("BoxMv<0: incorrect results are possible\n");
  if (s->BoxPs > GridNx * ds) printf("BoxPs>gridX: incorrect results are possible\n");
  if (s->BoxPa > GridNy * da) printf("BoxPa>gridY: incorrect results are possible\n");
  if (s->BoxPv > GridNz * dv) printf("BoxPv>gridZ: incorrect results are possible\n");
}

int calcStep(void) {
  if (!g_run.s) {
    fprintf(stderr, "calcStep: no simulation (call _main first)\n");
    return -1;
  }
  if (g_run.iStep == 0) printf("Starting...\n");
  double ms = 0;
  if (dtm_time_steps(g_run.s, Ntime, &ms) < 0) {
    fprintf(stderr, "calcStep: %s\n", dtm_last_error());
    return -1;
  }
  // Sensor::write once per calcStep (diamond.cu:473-475, sensor.h:81)
  if (!g_run.sensors.empty()) {
    std::vector<float> last((size_t)dtm_probe_count(g_run.s));
    if (dtm_probe_last(g_run.s, last.data()) > 0)
      for (auto& sf : g_run.sensors) {
        fprintf(sf.f, "%g\n", last[sf.probe]);
        fflush(sf.f);
      }
  }
  // SeismoDrops::drop after the step (diamond.cu:469-470, drop.cu:75-107); the
  // rate line counts its time as the reference's calcTime does
  const auto d0 = std::chrono::steady_clock::now();
  for (int c = 0; c < 6; c++) {
    if (!((DropData >> c) & 1)) continue;
    static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    char fname[1024];
    snprintf(fname, sizeof(fname), "%s/rag%s_%05d.arr", g_run.drop_dir.c_str(), names[c], g_run.iStep);
    if (dtm_write_arr(g_run.s, c, fname) < 0) {
      fprintf(stderr, "calcStep: %s\n", dtm_last_error());
      return -1;
    }
  }
  const double drop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d0).count();
  const double calc_ms = ms + drop_ms;
  const unsigned long long yee_cells = (unsigned long long)Ntime * GridNx * GridNy * GridNz;
  const int isTFSF = (g_run.iStep + 1) * Ntime * dt < shotpoint.tStop;
  printf("Step %d /node %d/ subnode %d/: Time %9.09f ms |drop %3.03f%% ||rate %9.09f GYee_cells/sec |total grid "
         "%dx%dx%d=%llu cells | isTFSF=%d\n",
         g_run.iStep, 0, 0, calc_ms, 100.0 * drop_ms / calc_ms, 1.e-9 * yee_cells / (calc_ms * 1.e-3), GridNx, GridNy, GridNz,
         (unsigned long long)GridNx * GridNy * GridNz, isTFSF);  // diamond.cu:483-485
  g_run.iStep++;
  fflush(stdout);
  return 0;
}

int _main(int argc, char** argv) {
  bool test_only = false;
  std::string drop_dir = ".";
  argv++;
  argc--;
  std::vector<int> extra;
  while (argc > 0 && strncmp(*argv, "--", 2) == 0) {
    if (strncmp(*argv, "--help", 6) == 0) return print_help();
    if (strcmp(*argv, "--test") == 0) {
      test_only = true;
      argv++;
      argc--;
      continue;
    }
    if (argc < 2) {
      printf("Illegal parameters' syntax notation\n");
      return print_help();
    }
    if (strcmp(*argv, "--drop_dir") == 0) drop_dir = argv[1];
    else if (strcmp(*argv, "--swap_dir") == 0) { /* no out-of-core swap in this build */ }
    else if (strcmp(*argv, "--sensor") == 0) {
      float v[3] = {0, 0, 0};
      char* p = argv[1];
      for (int i = 0; i < 3; i++) {
        v[i] = strtof(p, &p);
        if (*p) p++;
      }
      extra.push_back((int)v[0]);
      extra.push_back((int)v[1]);
      extra.push_back((int)v[2]);
    } else if (!strcmp(*argv, "--box") || !strcmp(*argv, "--mesh") || !strcmp(*argv, "--Dmesh") ||
               !strcmp(*argv, "--zoom") || !strcmp(*argv, "--step") || !strcmp(*argv, "--bkgr_col") ||
               !strcmp(*argv, "--mesh_col") || !strcmp(*argv, "--box_col")) { /* GL options */ }
    else {
      printf("Illegal parameters' syntax notation\n");
      return print_help();
    }
    printf("par: %s; vals: %s\n", argv[0], argv[1]);
    argv += 2;
    argc -= 2;
  }
  if (!test_only) {
    printf("With GL: the OpenGL viewer is not part of this build; use --test\n");
    return -1;
  }
  printf("No GL\n");
  mkdir(drop_dir.c_str(), 0700);
  if (DropData) {  // SeismoDrops::init (drop.cu:26-38): remove the snapshots of an earlier run
    static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    for (int it = 0;; it++) {
      int gone = 0;
      for (int c = 0; c < 6; c++) {
        char fname[1024];
        snprintf(fname, sizeof(fname), "%s/rag%s_%05d.arr", drop_dir.c_str(), names[c], it);
        if (remove(fname) != 0) gone++;
      }
      if (gone == 6) break;
    }
  }
  dtm_config c;
  dtm_config_defaults(&c, GridNx, GridNy, GridNz);
  c.dx = ds;
  c.dy = da;
  c.dz = dv;
  c.dt = dt;
  c.npml[0] = c.npml[1] = Npmlx;
  c.npml[2] = c.npml[3] = Npmly;
  // no layer at the low-z face, as in the reference (k2z = 1 for the low
  // half of its z table, DTmxw.cu:262): that face lies in the silver
  c.npml[4] = 0;
  c.npml[5] = Npmlz;
  c.pml_vmax = shotpoint.V_max > 0 ? shotpoint.V_max : 1.0;
  c.pml_kind = DTM_PML_SPLIT;  // genAlg.py:117-148
  c.pml_profile = DTM_PML_PROFILE_REFERENCE;
  c.ntime = Ntime;
  set_materials(c);
  // DipoleTFSF around the shot point, scattered field inside (signal.hpp:135-160)
  c.tfsf.kind = DTM_TFSF_DIPOLE;
  c.tfsf.waveform = DTM_WAVE_RAISED_COS;
  c.tfsf.sf_inside = 1;
  const float box[6] = {shotpoint.BoxMs, shotpoint.BoxPs, shotpoint.BoxMa, shotpoint.BoxPa, shotpoint.BoxMv,
                        shotpoint.BoxPv};
  for (int a = 0; a < 6; a++) c.tfsf.box[a] = box[a];
  c.tfsf.w = shotpoint.w;
  c.tfsf.kEnv = shotpoint.kEnv;
  c.tfsf.cSrc = shotpoint.cSrc;
  c.tfsf.Imp = shotpoint.Imp;
  c.tfsf.t0 = shotpoint.t0;
  c.tfsf.tw = shotpoint.tw;
  c.tfsf.amp = 1.0;
  c.tfsf.tStop = shotpoint.tStop;
  if (c.dt * sqrt(1 / (c.dx * c.dx) + 1 / (c.dy * c.dy) + 1 / (c.dz * c.dz)) > 6. / 7.) {
    printf("Error: Courant condition is not satisfied\n");
    return -1;
  }
  printf("Grid size: %dx%dx%d Yee_cells, TorreH=%d\n", GridNx, GridNy, GridNz, Ntime);
  if (g_run.s) dtm_destroy(g_run.s);
  for (auto& sf : g_run.sensors) fclose(sf.f);
  g_run = LegacyRun();
  g_run.drop_dir = drop_dir.empty() ? "." : drop_dir;
  if (dtm_create(&c, &g_run.s) < 0) {
    printf("Error: %s\n", dtm_last_error());
    return -1;
  }
  std::vector<uint8_t> m[3];
  build_geometry(m);
  if (dtm_set_materials(g_run.s, m[0].data(), m[1].data(), m[2].data()) < 0) {
    printf("Error: %s\n", dtm_last_error());
    return -1;
  }
  TFSFsrc_check(&shotpoint);
  // default sensors of init_index (DTmxw.cu:524-542)
  const double Xdip = (shotpoint.BoxMs + shotpoint.BoxPs) * 0.5, Ydip = (shotpoint.BoxMa + shotpoint.BoxPa) * 0.5;
  const int X0 = (int)lround(Xdip / ds), Y0 = (int)lround(Ydip / da), Z0 = (int)lround(0.150 / dv);
  for (double xsen = 0.0; xsen < GridNx * ds / 2; xsen += 1.0) {
    const int xcrd = X0 + (int)lround(xsen / ds);
    if (xcrd >= 0 && xcrd < GridNx && Y0 >= 0 && Y0 < GridNy && Z0 >= 0 && Z0 < GridNz) add_sensor_files(xcrd, Y0, Z0);
  }
  for (size_t q = 0; q + 2 < extra.size(); q += 3) add_sensor_files(extra[q], extra[q + 1], extra[q + 2]);
  printf("Full %d Big steps\n", Tsteps / Ntime);
  for (int i = 0; i < Tsteps / Ntime; i++)
    if (calcStep() < 0) return -1;
  // The sensor files stay open, as the reference's Sensor objects keep
  // theirs (sensor.h:15-18): a driver may go on calling calcStep.  The next
  // _main closes them.
  return 0;
}

}  // extern "C"

// Command-line driver over reconstruct() for the bin_data file layout: the
// same long options, defaults, input files and output files as the
// reference's driver (src/reconstruct/bin_data/XT_Main.c:47-215,
// XT_Main.h), with one process and no MPI.
//
//   tb_bin_data --proj_rows R --proj_cols C --proj_num P --recon_num N
//               --vox_wid W --rot_center X --sig_s S --sig_t T --c_s A --c_t B
//               --convg_thresh K --remove_rings 0|1|2|3 [--quad_convex]
//               [--huber_delta D] [--huber_T T] [--restart]
//
// Inputs in the working directory: projections.bin and weights.bin
// (float32, [proj_num][proj_rows][proj_cols]), proj_angles.bin and
// proj_times.bin (proj_num float32 each), recon_times.bin (recon_num + 1
// float32).  Outputs are the files reconstruct() writes there.
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "mbir4d.h"

namespace {

// generalized-Huber defaults of the reference driver (bin_data/XT_Main.h)
constexpr float kHuberDelta = 0.5f;
constexpr float kHuberT = 4.0f;

struct Args {
  int32_t proj_rows = 0, proj_cols = 0, proj_num = 0, recon_num = 0;
  float vox_wid = 0, rot_center = 0, sig_s = 0, sig_t = 0, c_s = 0, c_t = 0;
  float convg_thresh = 0;
  int32_t remove_rings = 0, quad_convex = 0;
  float huber_delta = kHuberDelta, huber_T = kHuberT;
  uint8_t restart = 0;
};

// The reference's option table (XT_Main.c:137-176): long names, short codes
// a..p, --quad_convex and --restart without a value.
int parse_args(int argc, char **argv, Args &a, FILE *log) {
  static const option opts[] = {
      {"proj_rows", required_argument, 0, 'a'},    {"proj_cols", required_argument, 0, 'b'},
      {"proj_num", required_argument, 0, 'c'},     {"recon_num", required_argument, 0, 'd'},
      {"vox_wid", required_argument, 0, 'e'},      {"rot_center", required_argument, 0, 'f'},
      {"sig_s", required_argument, 0, 'g'},        {"sig_t", required_argument, 0, 'h'},
      {"c_s", required_argument, 0, 'i'},          {"c_t", required_argument, 0, 'j'},
      {"convg_thresh", required_argument, 0, 'k'}, {"remove_rings", required_argument, 0, 'l'},
      {"quad_convex", no_argument, 0, 'm'},        {"huber_delta", required_argument, 0, 'o'},
      {"huber_T", required_argument, 0, 'p'},      {"restart", no_argument, 0, 'n'},
      {0, 0, 0, 0}};
  int bad = 0;
  for (;;) {
    int idx = 0;
    const int c = getopt_long(argc, argv, "a:b:c:d:e:f:g:h:i:j:k:l:mno:p:", opts, &idx);
    if (c == -1) break;
    switch (c) {
      case 'a': a.proj_rows = (int32_t)atoi(optarg); break;
      case 'b': a.proj_cols = (int32_t)atoi(optarg); break;
      case 'c': a.proj_num = (int32_t)atoi(optarg); break;
      case 'd': a.recon_num = (int32_t)atoi(optarg); break;
      case 'e': a.vox_wid = (float)atof(optarg); break;
      case 'f': a.rot_center = (float)atof(optarg); break;
      case 'g': a.sig_s = (float)atof(optarg); break;
      case 'h': a.sig_t = (float)atof(optarg); break;
      case 'i': a.c_s = (float)atof(optarg); break;
      case 'j': a.c_t = (float)atof(optarg); break;
      case 'k': a.convg_thresh = (float)atof(optarg); break;
      case 'l': a.remove_rings = (int32_t)atoi(optarg); break;
      case 'm': a.quad_convex = 1; break;
      case 'o': a.huber_delta = (float)atof(optarg); break;
      case 'p': a.huber_T = (float)atof(optarg); break;
      case 'n': a.restart = 1; break;
      default:
        fprintf(log, "ERROR: read_command_line_args: Cannot recognize an argument\n");
        bad = 1;
        break;
    }
  }
  if (argc - optind > 0) {
    fprintf(log, "ERROR: read_command_line_args: Argument list has an error\n");
    bad = 1;
  }
  if (a.proj_rows <= 0 || a.proj_cols <= 0 || a.proj_num <= 0 || a.recon_num <= 0) {
    fprintf(log, "ERROR: read_command_line_args: proj_rows, proj_cols, proj_num and recon_num "
                 "must be positive\n");
    bad = 1;
  }
  return bad;
}

// n floats of <name>.bin into a malloc'd buffer (reconstruct() frees its
// inputs, as the reference's does); nullptr and a message on a short file.
float *read_floats(const char *name, size_t n, FILE *log) {
  char path[256];
  snprintf(path, sizeof(path), "%s.bin", name);
  float *buf = (float *)malloc(n * sizeof(float));
  FILE *fp = fopen(path, "rb");
  size_t got = 0;
  if (buf && fp) got = fread(buf, sizeof(float), n, fp);
  if (fp) fclose(fp);
  if (!buf || got != n) {
    fprintf(log,
            "ERROR: read_BinFile: Read %zu number of elements from the file %s. However, "
            "required number of elements is %zu.\n",
            got, path, n);
    free(buf);
    return nullptr;
  }
  return buf;
}

}  // namespace

int main(int argc, char **argv) {
  FILE *log = stdout;
  Args a;
  if (parse_args(argc, argv, a, log)) return 1;
  fprintf(log,
          "main: Number of nodes is 1 and command line input argument values are proj_rows = "
          "%d, proj_cols = %d, proj_num = %d, recon_num = %d, vox_wid = %f, rot_center = %f, "
          "sig_s = %f, sig_t = %f, c_s = %f, c_t = %f, convg_thresh = %f, remove_rings = %d, "
          "quad_convex = %d, huber_delta = %f, huber_T = %f, restart = %d\n",
          a.proj_rows, a.proj_cols, a.proj_num, a.recon_num, a.vox_wid, a.rot_center, a.sig_s,
          a.sig_t, a.c_s, a.c_t, a.convg_thresh, a.remove_rings, a.quad_convex, a.huber_delta,
          a.huber_T, a.restart);
  fprintf(log, "main: Reading data ....\n");
  const size_t nsino = (size_t)a.proj_num * a.proj_rows * a.proj_cols;
  float *proj = read_floats("projections", nsino, log);
  float *wts = proj ? read_floats("weights", nsino, log) : nullptr;
  float *ang = wts ? read_floats("proj_angles", (size_t)a.proj_num, log) : nullptr;
  float *pt = ang ? read_floats("proj_times", (size_t)a.proj_num, log) : nullptr;
  float *rt = pt ? read_floats("recon_times", (size_t)a.recon_num + 1, log) : nullptr;
  if (!rt) {
    free(proj);
    free(wts);
    free(ang);
    free(pt);
    return 1;
  }
  fprintf(log, "main: Reconstructing the data ....\n");
  fflush(log);
  float *object = nullptr;
  const int rc = reconstruct(&object, proj, wts, ang, pt, rt, a.proj_rows, a.proj_cols,
                             a.proj_num, a.recon_num, a.vox_wid, a.rot_center, a.sig_s, a.sig_t,
                             a.c_s, a.c_t, a.convg_thresh, a.remove_rings, a.quad_convex,
                             a.huber_delta, a.huber_T, a.restart, log);
  free(object);
  fflush(log);
  return rc == 0 ? 0 : 1;
}

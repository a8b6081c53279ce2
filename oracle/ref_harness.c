/*
 * ref_harness.c -- drives the MPI-free part of the reference library
 * (/root/reference/src/MBIR_4D, compiled unmodified by oracle/Makefile) to
 * emit known-answer vectors for the oracle.  TEST INFRASTRUCTURE ONLY; this
 * file is ours, only the linked objects are the reference's.
 *
 * Modes (all binary outputs are little-endian raw arrays):
 *   rng SEED N OUT                       random2() after srandom2(SEED)
 *   hr ANGLES NP DXY DR OFFR BEAM OUT    DetectorResponseProfile -> f32[NP][65]
 *   acol HR ANGLES NP NX NY X0 Y0 DXY R0 RMAX DR NR OFFR OUT
 *        calcAMatrixColumnforAngle for every (view,row,col):
 *        per entry: i32 count, f64 values[8], i32 index[8]
 *   vsc MAG NT NB NY NX USEL SEED OUT
 *        VSC_based_Voxel_Line_Select after srandom2(SEED):
 *        per (t,b): i32 n, i32 x[n], i32 y[n]; then f64 next random2()
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include "XT_Structures.h"
#include "XT_Profile.h"
#include "XT_AMatrix.h"
#include "XT_NHICD.h"
#include "allocate.h"
#include "randlib.h"

static void *slurp(const char *path, size_t *bytes) {
  FILE *fp = fopen(path, "rb");
  if (!fp) { perror(path); exit(2); }
  fseek(fp, 0, SEEK_END);
  long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  void *buf = malloc(n ? n : 1);
  if (fread(buf, 1, n, fp) != (size_t)n) { perror(path); exit(2); }
  fclose(fp);
  if (bytes) *bytes = n;
  return buf;
}

static FILE *out_open(const char *path) {
  FILE *fp = fopen(path, "wb");
  if (!fp) { perror(path); exit(2); }
  return fp;
}

int main(int argc, char **argv) {
  if (argc < 2) return 1;
  FILE *devnull = fopen("/dev/null", "w");
  Sinogram S;
  ScannedObject O;
  TomoInputs T;
  memset(&S, 0, sizeof(S));
  memset(&O, 0, sizeof(O));
  memset(&T, 0, sizeof(T));
  T.debug_file_ptr = devnull;
  T.Write2Tiff = 0;
  T.node_rank = 1; /* silences the rank-0 debug prints */

  if (!strcmp(argv[1], "rng") && argc == 5) {
    uint32_t seed = (uint32_t)strtoul(argv[2], 0, 10);
    long n = atol(argv[3]);
    FILE *fp = out_open(argv[4]);
    srandom2(seed);
    for (long i = 0; i < n; i++) {
      double r = random2();
      fwrite(&r, sizeof(double), 1, fp);
    }
    fclose(fp);
    return 0;
  }
  if (!strcmp(argv[1], "hr") && argc == 9) {
    float *angles = (float *)slurp(argv[2], 0);
    S.N_p = atoi(argv[3]);
    O.delta_xy = atof(argv[4]);
    S.delta_r = atof(argv[5]);
    S.OffsetR = atof(argv[6]);
    O.BeamWidth = atof(argv[7]);
    O.delta_z = O.delta_xy;
    S.delta_t = S.delta_r;
    S.OffsetT = (O.delta_z / 2 + S.delta_t / 2) / DETECTOR_RESPONSE_BINS;
    S.ViewPtr = angles;
    Real_arr_t **H_r = (Real_arr_t **)multialloc(sizeof(Real_arr_t), 2, S.N_p, DETECTOR_RESPONSE_BINS + 1);
    Real_arr_t *H_t = (Real_arr_t *)get_spc(DETECTOR_RESPONSE_BINS + 1, sizeof(Real_arr_t));
    DetectorResponseProfile(H_r, H_t, &S, &O, &T);
    FILE *fp = out_open(argv[8]);
    fwrite(&H_r[0][0], sizeof(Real_arr_t), (size_t)S.N_p * (DETECTOR_RESPONSE_BINS + 1), fp);
    fclose(fp);
    return 0;
  }
  if (!strcmp(argv[1], "acol") && argc == 16) {
    size_t hb;
    float *hr = (float *)slurp(argv[2], &hb);
    float *angles = (float *)slurp(argv[3], 0);
    S.N_p = atoi(argv[4]);
    O.N_x = atoi(argv[5]);
    O.N_y = atoi(argv[6]);
    O.x0 = atof(argv[7]);
    O.y0 = atof(argv[8]);
    O.delta_xy = atof(argv[9]);
    S.R0 = atof(argv[10]);
    S.RMax = atof(argv[11]);
    S.delta_r = atof(argv[12]);
    S.N_r = atoi(argv[13]);
    S.OffsetR = atof(argv[14]);
    Real_arr_t **H_r = (Real_arr_t **)multialloc(sizeof(Real_arr_t), 2, S.N_p, DETECTOR_RESPONSE_BINS + 1);
    memcpy(&H_r[0][0], hr, (size_t)S.N_p * (DETECTOR_RESPONSE_BINS + 1) * sizeof(float));
    S.cosine = (Real_t *)get_spc(S.N_p, sizeof(Real_t));
    S.sine = (Real_t *)get_spc(S.N_p, sizeof(Real_t));
    for (int i = 0; i < S.N_p; i++) { /* calculateSinCos, XT_Init.c:193-204 */
      S.cosine[i] = cos(angles[i]);
      S.sine[i] = sin(angles[i]);
    }
    AMatrixCol col;
    col.values = (Real_t *)get_spc(64, sizeof(Real_t));
    col.index = (int32_t *)get_spc(64, sizeof(int32_t));
    FILE *fp = out_open(argv[15]);
    for (int v = 0; v < S.N_p; v++)
      for (int r = 0; r < O.N_y; r++)
        for (int c = 0; c < O.N_x; c++) {
          calcAMatrixColumnforAngle(&S, &O, H_r, &col, r, c, v);
          int32_t n = col.count;
          double vals[8] = {0};
          int32_t idx[8] = {0};
          for (int q = 0; q < n && q < 8; q++) {
            vals[q] = col.values[q];
            idx[q] = col.index[q];
          }
          fwrite(&n, sizeof(int32_t), 1, fp);
          fwrite(vals, sizeof(double), 8, fp);
          fwrite(idx, sizeof(int32_t), 8, fp);
        }
    fclose(fp);
    return 0;
  }
  if (!strcmp(argv[1], "vsc") && argc == 10) {
    float *mag = (float *)slurp(argv[2], 0);
    O.N_time = atoi(argv[3]);
    T.num_z_blocks = atoi(argv[4]);
    O.N_y = atoi(argv[5]);
    O.N_x = atoi(argv[6]);
    int usel = atoi(argv[7]);
    uint32_t seed = (uint32_t)strtoul(argv[8], 0, 10);
    O.NHICD_Iterations = 10;
    int nxy = O.N_x * O.N_y;
    Real_arr_t ****M = (Real_arr_t ****)multialloc(sizeof(Real_arr_t), 4, O.N_time, T.num_z_blocks, O.N_y, O.N_x);
    memcpy(&M[0][0][0][0], mag, (size_t)O.N_time * T.num_z_blocks * nxy * sizeof(float));
    T.x_NHICD_select = (int32_t ***)multialloc(sizeof(int32_t), 3, O.N_time, T.num_z_blocks, nxy);
    T.y_NHICD_select = (int32_t ***)multialloc(sizeof(int32_t), 3, O.N_time, T.num_z_blocks, nxy);
    T.UpdateSelectNum = (int32_t **)multialloc(sizeof(int32_t), 2, O.N_time, T.num_z_blocks);
    T.NHICDSelectNum = (int32_t **)multialloc(sizeof(int32_t), 2, O.N_time, T.num_z_blocks);
    for (int i = 0; i < O.N_time; i++)
      for (int b = 0; b < T.num_z_blocks; b++) T.UpdateSelectNum[i][b] = usel;
    srandom2(seed);
    VSC_based_Voxel_Line_Select(&O, &T, M);
    FILE *fp = out_open(argv[9]);
    for (int i = 0; i < O.N_time; i++)
      for (int b = 0; b < T.num_z_blocks; b++) {
        int32_t n = T.NHICDSelectNum[i][b];
        fwrite(&n, sizeof(int32_t), 1, fp);
        fwrite(T.x_NHICD_select[i][b], sizeof(int32_t), n, fp);
        fwrite(T.y_NHICD_select[i][b], sizeof(int32_t), n, fp);
      }
    double nxt = random2();
    fwrite(&nxt, sizeof(double), 1, fp);
    fclose(fp);
    return 0;
  }
  fprintf(stderr, "ref_harness: bad arguments\n");
  return 1;
}

// reconstruct(): the MBIR_4D entry point (src/include/mbir4d.h:34, defined
// at src/MBIR_4D/XT_Engine.c:81-203) on top of the GPU stage API.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "mbir4d.h"
#include "timbir_dev.h"
#include "tb_common.h"

namespace {

// XT_Debug.h:8-32 message style
void log_info(FILE *fp, bool on, const char *fmt, ...) {
  if (!fp || !on) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(fp, "[INFO] (%s:%d) ", __FILE__, __LINE__);
  vfprintf(fp, fmt, ap);
  va_end(ap);
}

void log_err(FILE *fp, const char *fmt, ...) {
  if (!fp) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(fp, "[ERROR] (%s:%d: errno: None) ", __FILE__, __LINE__);
  vfprintf(fp, fmt, ap);
  va_end(ap);
}

int write_bin(const char *name, const void *data, size_t bytes) {
  FILE *fp = fopen(name, "wb");
  if (!fp) return -1;
  size_t w = fwrite(data, 1, bytes, fp);
  fclose(fp);
  return w == bytes ? 0 : -1;
}

int append_bin(const char *name, const void *data, size_t bytes) {
  FILE *fp = fopen(name, "ab");
  if (!fp) return -1;
  size_t w = fwrite(data, 1, bytes, fp);
  fclose(fp);
  return w == bytes ? 0 : -1;
}

int read_bin(const char *name, void *data, size_t bytes) {
  FILE *fp = fopen(name, "rb");
  if (!fp) return -1;
  size_t r = fread(data, 1, bytes, fp);
  fclose(fp);
  return r == bytes ? 0 : -1;
}

// write_ObjectProjOff2TiffBinPerIter (XT_IOMisc.c:238-271) .bin part plus the
// stage-end writes of ICD_BackProject (XT_ICD_update.c:1205-1213)
int write_stage_outputs(tb_stage *st, int stage, const std::vector<float> &obj,
                        const std::vector<float> &mag, const std::vector<float> &off,
                        double var_est, int N_time, size_t plane) {
  char name[64];
  for (int f = 0; f < N_time; f++) {
    snprintf(name, sizeof(name), "object_time_%d.bin", f);
    if (write_bin(name, obj.data() + (size_t)f * plane, plane * sizeof(float))) return -1;
  }
  if (write_bin("proj_offset.bin", off.data(), off.size() * sizeof(float))) return -1;
  if (write_bin("variance_estimate.bin", &var_est, sizeof(double))) return -1;
  if (write_bin("mag_update.bin", mag.data(), mag.size() * sizeof(float))) return -1;
  if (write_bin("status.bin", &stage, sizeof(int32_t))) return -1;
  (void)st;
  return 0;
}

}  // namespace

using tb::kMaxMultres;
using tb::kMinRowsPerNode;
using tb::kMinXYRes;

// The multi-resolution schedule of reconstruct() with one node
// (XT_Engine.c:111-141): stages coarse to fine.
extern "C" int32_t tb_multires_schedule(int32_t proj_rows, int32_t proj_cols, int32_t *mult_xy,
                                        int32_t *mult_z) {
  if (proj_cols <= 0 || proj_rows <= 0) return -1;
  int nres = (int)(log(((float)proj_cols) / kMinXYRes) / log(2.0) + 1);
  if (nres > kMaxMultres) nres = kMaxMultres;
  if (nres < 2) nres = 2;
  int mxy[kMaxMultres], mz[kMaxMultres];
  mxy[0] = 1;
  mz[0] = 1;
  for (int i = 1; i < nres; i++) {
    if (proj_cols % (mxy[i - 1] * 2) == 0) {
      mxy[i] = mxy[i - 1] * 2;
      if (proj_rows % (mz[i - 1] * 2) == 0 && proj_rows / (mz[i - 1] * 2) >= kMinRowsPerNode)
        mz[i] = mz[i - 1] * 2;
      else
        mz[i] = mz[i - 1];
    } else {
      nres = i;
      break;
    }
  }
  for (int i = 0; i < nres; i++) {
    if (mult_xy) mult_xy[i] = mxy[nres - 1 - i];
    if (mult_z) mult_z[i] = mz[nres - 1 - i];
  }
  return nres;
}

extern "C" int reconstruct(float **object, float *projections, float *weights,
                           float *proj_angles, float *proj_times, float *recon_times,
                           int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                           int32_t recon_num, float vox_wid, float rot_center, float sig_s,
                           float sig_t, float c_s, float c_t, float convg_thresh,
                           int32_t remove_rings, int32_t quad_convex, float huber_delta,
                           float huber_T, uint8_t restart, FILE *debug_msg_ptr) {
  using namespace tb;
  tb_options opt;
  tb_get_options(&opt);
  const bool verbose = opt.verbose != 0;
  FILE *dbg = debug_msg_ptr;
  log_info(dbg, verbose, "Reconstructing the data ....\n");
  const time_t start = time(NULL);
  // XT_Engine.c:101-104 with one node
  if (proj_rows < kMinProjRows) {
    log_err(dbg, "The minimum number of projection rows should be %d.\n", kMinProjRows);
    return -1;
  }
  if (proj_cols % 2 != 0) {
    log_err(dbg, "The number of projection columns should be even.\n");
    return -1;
  }
  if (proj_rows % 2 != 0 || proj_rows < kMinRowsPerNode) {
    log_err(dbg,
            "The number of projection rows divided by the number of nodes should be an even "
            "number greater than or equal to %d.\n",
            kMinRowsPerNode);
    return -1;
  }
  // multi-resolution schedule, XT_Engine.c:111-141
  int32_t sxy[kMaxMultres], sz[kMaxMultres];
  const int nres = tb_multires_schedule(proj_rows, proj_cols, sxy, sz);
  log_info(dbg, verbose, "Number of multi-resolution stages is %d.\n", nres);

  // restart (XT_Engine.c:143-157): the loop starts again at the stage
  // status.bin records, from that stage's own files read without
  // interpolation (initObject's initICD = 1 path, XT_ICD_update.c:558-571)
  int first = 0;
  bool resume = false;
  std::vector<float> prev_obj, prev_mag, prev_off;
  double var_est = 1.0;  // VAR_PARAM_INIT
  int prev_nz = 0;
  if (restart == 1) {
    int32_t last = -1;
    if (read_bin("status.bin", &last, sizeof(last)) || last < 0 || last >= nres) {
      log_err(dbg, "Cannot read a valid status.bin for restart.\n");
      return -1;
    }
    first = last;
    resume = true;
    {
      // the files of the last completed stage: the next stage starts from
      // them, or they are the result when every stage is done
      const int Nx = proj_cols / sxy[last], Nz = proj_rows / sz[last];
      const size_t plane = (size_t)Nz * Nx * Nx;
      prev_obj.resize(plane * recon_num);
      prev_mag.resize(plane * recon_num);
      prev_off.resize((size_t)proj_cols * proj_rows);
      char name[64];
      for (int f = 0; f < recon_num; f++) {
        snprintf(name, sizeof(name), "object_time_%d.bin", f);
        if (read_bin(name, prev_obj.data() + plane * f, plane * sizeof(float))) {
          log_err(dbg, "Error in reading file %s.\n", name);
          return -1;
        }
      }
      if (read_bin("mag_update.bin", prev_mag.data(), prev_mag.size() * sizeof(float)) ||
          read_bin("proj_offset.bin", prev_off.data(), prev_off.size() * sizeof(float)) ||
          read_bin("variance_estimate.bin", &var_est, sizeof(double))) {
        log_err(dbg, "Error in reading the restart files.\n");
        return -1;
      }
      prev_nz = Nz;
    }
  }

  // projs_runtime.bin / weights_runtime.bin: the transposed scan the
  // reference re-reads at every stage (XT_Init.c:298-315)
  // the scan is kept here for the later stages, as the reference keeps it in
  // projs_runtime.bin / weights_runtime.bin
  const size_t nraw = (size_t)proj_num * proj_rows * proj_cols;
  std::vector<float> raw_p(projections, projections + nraw), raw_w(weights, weights + nraw);
  bool freed_inputs = false;
  std::vector<float> final_obj;
  int final_nz = 0;
  tb_stage *prev_st = nullptr;  // the previous stage, kept in HBM for the handoff
  for (int s = first; s < nres; s++) {
    log_info(dbg, verbose,
             "Running multi-resolution stage %d with x-y voxel scale = %d and z voxel scale = "
             "%d.\n",
             s, sxy[s], sz[s]);
    tb_stage *st = nullptr;
    if (tb_stage_create(&st, raw_p.data(), raw_w.data(), proj_angles, proj_times, recon_times,
                        proj_rows, proj_cols, proj_num, recon_num, vox_wid, rot_center, sig_s,
                        sig_t, c_s, c_t, convg_thresh, remove_rings, quad_convex, huber_delta,
                        huber_T, sxy[s], sz[s], s, s ? sz[s - 1] : 0, var_est)) {
      log_err(dbg, "%s\n", tb_last_error());
      tb_stage_destroy(prev_st);
      return -1;
    }
    int32_t ints[16];
    double dbls[64];
    tb_stage_info(st, ints, dbls);
    const int N_x = ints[3], N_y = ints[4], N_z = ints[5];
    const size_t plane = (size_t)N_z * N_y * N_x;
    if (s == first && opt.write_files) {
      std::vector<float> tr((size_t)proj_num * proj_rows * proj_cols);
      for (int i = 0; i < proj_num; i++)
        for (int j = 0; j < proj_cols; j++)
          for (int k = 0; k < proj_rows; k++)
            tr[((size_t)i * proj_cols + j) * proj_rows + k] =
                raw_p[((size_t)i * proj_rows + k) * proj_cols + j];
      write_bin("projs_runtime.bin", tr.data(), tr.size() * sizeof(float));
      for (int i = 0; i < proj_num; i++)
        for (int j = 0; j < proj_cols; j++)
          for (int k = 0; k < proj_rows; k++)
            tr[((size_t)i * proj_cols + j) * proj_rows + k] =
                raw_w[((size_t)i * proj_rows + k) * proj_cols + j];
      write_bin("weights_runtime.bin", tr.data(), tr.size() * sizeof(float));
    }
    if (!freed_inputs) {
      // the reference frees the caller's scan after the first stage set-up
      // (XT_Engine.c:160-164); later stages keep using the uploaded copy
      free(projections);
      free(weights);
      projections = nullptr;
      weights = nullptr;
      freed_inputs = true;
    }
    int rc = 0;
    if (prev_st) {
      // object and map upsampled from the previous stage in HBM
      rc = tb_init_from_stage(st, prev_st);
      tb_stage_destroy(prev_st);
      prev_st = nullptr;
      if (!rc && !prev_off.empty()) rc = tb_set_offset(st, prev_off.data());
    } else if (resume && s == first) {
      // restart: this stage again, from its own files as they are
      rc = tb_set_object(st, prev_obj.data());
      if (!rc && s > 0) rc = tb_set_magmap(st, prev_mag.data());
      if (!rc) rc = tb_set_offset(st, prev_off.data());
    } else if (opt.init_fbp) {
      // stage 0 from the filtered back-projection instead of zero
      rc = tb_init_fbp(st);
    }
    int iters = 0;
    std::vector<double> costs;
    if (!rc) {
      if (opt.compute_cost) costs.assign((opt.max_iterations > 0 ? opt.max_iterations : kMaxIters) + 1, 0);
      rc = tb_stage_run(st, 0, opt.compute_cost ? costs.data() : nullptr, &iters);
    }
    if (rc) {
      log_err(dbg, "Reconstruction failed! %s\n", tb_last_error());
      tb_stage_destroy(st);
      return -1;
    }
    if (opt.compute_cost && opt.write_files) {
      write_bin("cost.bin", costs.data(), sizeof(double));
      for (int i = 1; i <= iters; i++) append_bin("cost.bin", &costs[i], sizeof(double));
    }
    // host copies only for the stage's files and for the final result
    const bool last = s == nres - 1;
    const bool host = opt.write_files || last;
    std::vector<float> obj(host ? plane * recon_num : 0), mag(host ? plane * recon_num : 0),
        off((size_t)proj_cols * proj_rows);
    rc = tb_get_offset(st, off.data());
    if (!rc && host) rc = tb_get_object(st, obj.data()) || tb_get_magmap(st, mag.data());
    var_est = tb_get_var_est(st);
    if (last) tb_stage_destroy(st);
    else prev_st = st;
    if (rc) {
      log_err(dbg, "%s\n", tb_last_error());
      tb_stage_destroy(prev_st);
      return -1;
    }
    log_info(dbg, verbose, "Time elapsed is %f minutes.\n", difftime(time(NULL), start) / 60.0);
    if (opt.write_files &&
        write_stage_outputs(nullptr, s, obj, mag, off, var_est, recon_num, plane)) {
      log_err(dbg, "Cannot write the stage outputs.\n");
      return -1;
    }
    log_info(dbg, verbose, "Completed multi-resolution stage %d after %d iterations.\n", s,
             iters);
    prev_obj.swap(obj);
    prev_mag.swap(mag);
    prev_off.swap(off);
    prev_nz = N_z;
    final_nz = N_z;
  }
  if (!freed_inputs) {
    free(projections);
    free(weights);
  }
  free(proj_angles);
  free(proj_times);
  free(recon_times);
  if (object) {
    const size_t n = prev_obj.size();
    float *o = (float *)malloc(n * sizeof(float));
    if (o) memcpy(o, prev_obj.data(), n * sizeof(float));
    *object = o;
  }
  (void)final_nz;
  log_info(dbg, verbose, "Exiting MBIR 4D\n");
  return 0;
}

/* CPU restatement of the reference's media binning MEX
 * (+MCmatlab/src/getDownsampledParamVals.c:54-202) — TEST INFRASTRUCTURE ONLY.
 * Only tests/ link it, as the checker of the GPU binning; the library never
 * calls it.
 *
 * [downsampledParamVals, binidx] = getDownsampledParamVals(paramVals, N2):
 * paramVals holds N1 parameter sets (mua, mus, g) as a 3 x N1 single matrix;
 * N2 sets are picked by farthest-point sampling in the range-normalised
 * parameter space of the rows that vary, starting from the set farthest from
 * the mean; binidx (1-based, uint8) names the picked set each input set is
 * nearest to.  The arithmetic follows the reference's C semantics:
 * single-precision running sums in input order (:64-83), normalisation and
 * mean in single precision (:92-104), squared distances accumulated as
 * float += (double)d * d (:101-103, :118-121, :154-157; pow(d, 2) of a float
 * is exact in double), and the selection of the single-threaded build
 * (:126-130: nThreads = 1 off Windows), whose running maximum keeps the index
 * of the previous pass (:164-166, calloc'd to 0 at :131) and whose winner
 * is compared against index 0 (:172-175).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

int oracle_downsample_params(const float *pv, int64_t N1, int64_t N2, float *out, uint8_t *binidx) {
  if (N1 < 1 || N2 < 1 || !pv || !out || !binidx) return -1;
  float mx[3], mn[3], sum[3];
  for (int r = 0; r < 3; ++r) mx[r] = mn[r] = sum[r] = pv[r];
  for (int64_t i = 1; i < N1; ++i) /* :73-83 */
    for (int r = 0; r < 3; ++r) {
      const float v = pv[3 * i + r];
      mx[r] = v > mx[r] ? v : mx[r];
      mn[r] = v < mn[r] ? v : mn[r];
      sum[r] += v;
    }
  int rowOf[3], rows = 0; /* :84-90: the varying rows, in order */
  for (int r = 0; r < 3; ++r) rowOf[r] = mx[r] != mn[r] ? rows++ : -1;
  float *norm = malloc(sizeof(float) * (size_t)(rows ? rows : 1) * (size_t)N1);
  float *d = malloc(sizeof(float) * (size_t)N1);
  float *ds = malloc(sizeof(float) * (size_t)(rows ? rows : 1) * (size_t)N2);
  if (!norm || !d || !ds) {
    free(norm), free(d), free(ds);
    return -2;
  }
  for (int64_t i = 0; i < N1; ++i) /* :92-96 */
    for (int r = 0; r < 3; ++r)
      if (rowOf[r] >= 0) norm[rows * i + rowOf[r]] = pv[3 * i + r] / (mx[r] - mn[r]);
  float mean[3];
  for (int r = 0; r < 3; ++r) mean[r] = rowOf[r] >= 0 ? sum[r] / (mx[r] - mn[r]) / (float)N1 : 0.f;
  int64_t c = 0; /* :98-105: farthest from the mean */
  for (int64_t i = 0; i < N1; ++i) {
    float s = 0.f;
    for (int r = 0; r < 3; ++r)
      if (rowOf[r] >= 0) {
        const float x = norm[rows * i + rowOf[r]] - mean[r];
        s = (float)((double)s + (double)x * (double)x);
      }
    d[i] = s;
    if (d[i] > d[c]) c = i;
  }
  for (int q = 0; q < rows; ++q) ds[q] = norm[rows * c + q]; /* :108-111 */
  int64_t best = 0; /* :113-124: distances to the first pick */
  for (int64_t i = 0; i < N1; ++i) {
    float s = 0.f;
    for (int q = 0; q < rows; ++q) {
      const float x = norm[rows * i + q] - ds[q];
      s = (float)((double)s + (double)x * (double)x);
    }
    d[i] = s;
    binidx[i] = 1;
    if (d[i] > d[best]) best = i;
  }
  c = best;
  int64_t m = 0; /* maxDistIdxs[0], calloc'd (:131) and kept between passes */
  for (int64_t j = 1; j < N2; ++j) {
    for (int q = 0; q < rows; ++q) ds[rows * j + q] = norm[rows * c + q]; /* :144-147 */
    for (int64_t k = 0; k < N1; ++k) { /* :153-167 */
      float s = 0.f;
      for (int q = 0; q < rows; ++q) {
        const float x = norm[rows * k + q] - ds[rows * j + q];
        s = (float)((double)s + (double)x * (double)x);
      }
      if (s < d[k]) {
        d[k] = s;
        binidx[k] = (uint8_t)(j + 1);
      }
      if (d[k] > d[m]) m = k;
    }
    c = d[m] > d[0] ? m : 0; /* :172-175 */
  }
  for (int64_t i = 0; i < N2; ++i) /* :182-186 */
    for (int r = 0; r < 3; ++r)
      out[3 * i + r] = rowOf[r] >= 0 ? ds[rows * i + rowOf[r]] * (mx[r] - mn[r]) : pv[r];
  free(norm), free(d), free(ds);
  return 0;
}

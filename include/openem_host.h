/* Host-side helpers of the MI355X FDTD stepper (no GPU needed).
 *
 * These build the compact tables the device library consumes.  They run on the CPU and are
 * exported from the same shared library as openem_hip.h, so a caller that has only the ABI
 * can build the material-class and pole-coefficient tables at native speed.
 */
#ifndef OPENEM_HOST_H
#define OPENEM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Factorize n rows of ncols 32-bit words given column-wise: word c of row r is
 * cols[c][r * stride[c]]; a NULL stride means 1 for every column.  A float64 column enters as
 * two word columns of stride 2 whose pointers are 4 bytes apart.  Distinct rows are numbered in
 * order of first appearance: inv[r] receives row r's number and uniq (row-major,
 * at least max_uniq * ncols words) the distinct rows.  *n_uniq receives their count, or -1 when
 * there are more than max_uniq (inv and uniq are then incomplete).  Returns OEM_OK or
 * OEM_EINVAL / OEM_ENOMEM.
 *
 * Replaces the numpy deduplications behind the reference's material-class LUT
 * (`openem/coeffs.py:174-208`, (ca, cb) pairs per component -> uint16 index) and its pole
 * coefficient LUT (`openem/dispersion_setup.py:37`, `_coef_lut`). */
int oem_host_factorize(int64_t n, int ncols, const uint32_t *const *cols,
                       const int32_t *stride, int32_t *inv, uint32_t *uniq, int64_t max_uniq,
                       int64_t *n_uniq);

#ifdef __cplusplus
}
#endif

#endif

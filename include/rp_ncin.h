/*
 * rp_ncin.h -- hydrodynamic input of librp.so: reads ROMS grid and history
 * files in the netCDF classic formats (CDF-1, CDF-2 with 64-bit offsets,
 * CDF-5 with 64-bit data), the format ROMS writes unless built for netCDF-4.
 * It stands in for the nf90_inq_dimid / nf90_inq_varid / nf90_get_var calls
 * of InitGrid (grid_module.f90:41-422) and updateHydro
 * (hydrodynamic_module.f90:112-760), because no netCDF library is present.
 * NETCDF4 (HDF5) files are rejected with RP_E_IO and a message saying so.
 *
 * Values are converted to double exactly as nf90_get_var converts them into
 * a DOUBLE PRECISION array (no scale_factor / add_offset, as the reference
 * applies none).  A record variable's slice comes back in file order: for
 * zeta(ocean_time, eta_rho, xi_rho) that is xi_rho fastest, i.e. the
 * Fortran array zeta(xi_rho, eta_rho) the reference fills.
 *
 * Host code only.  Every function returns 0 or a negative RP_E* code
 * (rp_abi.h); rp_ncin_error() gives the message.
 */
#ifndef RP_NCIN_H
#define RP_NCIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NF90_OPEN(path, NF90_NOWRITE): reads the header. */
int rp_ncin_open(void **f, const char *path);

/* NF90_INQ_DIMID + NF90_INQUIRE_DIMENSION(len): the length of a dimension
 * (the current number of records for the unlimited one). */
int rp_ncin_dim(void *f, const char *name, int64_t *len);

/* NF90_INQ_VARID + NF90_INQUIRE_VARIABLE: the shape in file order (slowest
 * first, the record dimension, when there is one, first and given as the
 * number of records), its rank (at most 8) and whether it is a record
 * variable.  Returns RP_E_ARG when there is no such variable. */
int rp_ncin_var(void *f, const char *name, int32_t *ndims, int64_t shape[8], int32_t *is_record);

/* NF90_GET_VAR: record rec (0-based) of a record variable, or the whole of
 * a fixed one (rec ignored), as double; count must equal the number of
 * values of that slice. */
int rp_ncin_get(void *f, const char *name, int64_t rec, double *out, int64_t count);

/* NF90_GET_ATT of a text attribute (var NULL: a global one) into buf,
 * NUL-terminated and cut to size; len gets the attribute's length. */
int rp_ncin_att_text(void *f, const char *var, const char *att, char *buf, int64_t size,
                     int64_t *len);

/* Message of the last failure (f NULL: of the last rp_ncin_open). */
const char *rp_ncin_error(void *f);

int rp_ncin_close(void *f);

#ifdef __cplusplus
}
#endif

#endif /* RP_NCIN_H */

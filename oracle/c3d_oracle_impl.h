/* Internal layout of the oracle grid, shared by the oracle's C files
 * (test infrastructure). */
#ifndef C3D_ORACLE_IMPL_H
#define C3D_ORACLE_IMPL_H

struct orc_grid {
  int nx, ny, nz, nxg, nyg, nzg, nxp, nzp, nym;
  double *x, *y, *z, *xg, *yg, *zg;
  double *kxx, *kzz;
  double *dyy_d, *dyy_l, *dyy_u; /* D(j), DL(j)=Dyy(j+1,j), DU(j)=Dyy(j,j+1) */
  double dx, dz, dxmin, dymin, dzmin, Lx, Ly, Lz, Lxp, Lzp;
  int n_uniform; /* uniform buffer points per wall (grid_type 2), else 1 */
};

#endif

/*
 * TEST INFRASTRUCTURE ONLY. C wrapper around the reference's own vendored HEALPix
 * (src/Healpix_2.15a/healpix_base.{h,cc}, compiled from /root/reference by oracle/Makefile into
 * oracle/_ref/). Used only by tests to pin the oracle's and the product's HEALPix NEST code.
 */
#include "src/Healpix_2.15a/healpix_base.h"

extern "C" {
void ref_hp_pix2ang_z_phi(int order, long long pix, double *z, double *phi) {
    Healpix_Base b(order, NEST);
    b.pix2ang_z_phi((int)pix, *z, *phi);
}
void ref_hp_nest2xyf(int order, long long pix, int *ix, int *iy, int *face) {
    Healpix_Base b(order, NEST);
    b.nest2xyf((int)pix, *ix, *iy, *face);
}
long long ref_hp_xyf2nest(int order, int ix, int iy, int face) {
    Healpix_Base b(order, NEST);
    return b.xyf2nest(ix, iy, face);
}
long long ref_hp_ang2pix_z_phi(int order, double z, double phi) {
    Healpix_Base b(order, NEST);
    return b.ang2pix_z_phi(z, phi);
}
long long ref_hp_npix(int order) {
    Healpix_Base b(order, NEST);
    return b.Npix();
}
}

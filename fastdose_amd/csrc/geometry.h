// geometry.h -- host geometry of libfastdose_hip (see geometry.cpp).
#pragma once

#include "fd_internal.h"

namespace fdi {

// Rotation matrix of a beam orientation (gantry, couch, collimator) composed in FP64 and rounded
// once to FP32, row-major.  to_pvcs = false gives PVCS -> BEV (geom.cu:36-44), true BEV -> PVCS
// (geom.cu:47-56).
void beam_matrix(double gantry, double couch, double coll, bool to_pvcs, float out[9]);

int calc_range(fd_beam& beam, const fd_volume& vol);
int beamlet_init(const fd_bundle& bb, int ix, int iy, const fd_volume& vol, fd_beam& out, float* fluence);
int beamlet_flags(const fd_bundle& bb, const uint8_t* mask, const fd_volume& vol, uint8_t* flags);
// Central beam of a bundle for the flag kernel (calc_range status returned; L = 0 on a miss).
int flag_beam_setup(const fd_bundle& bb, const fd_volume& vol, FlagBeamDev& d);
int ptv_isocenter(const uint8_t* mask, const fd_volume& vol, float iso[3]);

void fill_beamlet_dev(const fd_beam& b, uint64_t bev_off, BeamletDev& d);
// Everything but the rotation matrices Rf / Ri.
void fill_beamlet_dev_no_rotation(const fd_beam& b, uint64_t bev_off, BeamletDev& d);

}  // namespace fdi

/*
 * parameters.h -- the reference's step parameters, field for field
 * (src/parameters.h:21-58), so that callers of run_vbm3d compile unchanged.
 */
#ifndef VBM3D_AMD_PARAMETERS_H
#define VBM3D_AMD_PARAMETERS_H

struct Parameters
{
	/// side of the square patch, in pixels
	unsigned k;
	/// number of frames a patch spans (temporal extent)
	unsigned kt;
	/// spacing of the reference-patch grid, in pixels
	unsigned p;
	/// largest group: similar patches kept per reference patch
	unsigned N;
	/// distance above which a candidate leaves the group
	float tau;
	/// bonus subtracted from the distance at the reference position
	float d;
	/// frames searched after (and before) the reference frame
	unsigned Nf;
	/// side of the search window in the reference frame
	unsigned Ns;
	/// side of the predictive search windows in the other frames
	unsigned Npr;
	/// matches carried from one frame to the next
	unsigned Nb;
	/// code of the 2D transform of a patch (DCT or BIOR)
	unsigned T_2D;
	/// code of the 1D transform along the group (HADAMARD or HAAR)
	unsigned T_3D;
	/// hard threshold factor (only read by the hard-threshold step)
	float lambda3D = 0.;
	/// patches follow the optical flow across their kt frames
	bool mc = false;
	/// border of the spatial tiles of a multi-threaded run
	int n = 16;
};

#endif

/*
 * parameters.h -- the reference's step parameters, field for field
 * (src/parameters.h:21-58), so that callers of run_vbm3d compile unchanged.
 */
#ifndef VBM3D_AMD_PARAMETERS_H
#define VBM3D_AMD_PARAMETERS_H

struct Parameters
{
	/// Size of the patch (spatial)
	unsigned k;
	/// Size of the patch (temporal)
	unsigned kt;
	/// Step size
	unsigned p;
	/// Maximum number of similar patches
	unsigned N;
	/// Maximum distance threshold
	float tau;
	/// Correcting parameter in the distance computation
	float d;
	/// Number of frames forward (and backward) used during the search
	unsigned Nf;
	/// Size of the search region in the reference frame
	unsigned Ns;
	/// Size of the search region in the other frame
	unsigned Npr;
	/// Maximum number of matches kept for a frame
	unsigned Nb;
	/// Type of the 2D tranform
	unsigned T_2D;
	/// Type of the 1D tranform
	unsigned T_3D;
	/// Threshold if it's a hard thresholding step
	float lambda3D = 0.;
	/// Motion compensation for 3d patches
	bool mc = false;
	/// Border of the tile when using the multithreading
	int n = 16;
};

#endif

/*
 * 2d-treecodes-force.h (float) -- drop-in header of the reference's real=float build.
 * Same declaration as the reference's generated header
 * (libraries.Makefile:35-36 from drivers/treecode-force.h:21-31, realtype=float).
 * Implemented by libtreecodes2d_float.so (also installed as
 * 2d-treecodes_amd/float/lib2d-treecodes-force.so).
 * Output layout: the result for point (ix, iy) of block ib is at ix + 32 * (iy + 32 * ib).
 */
#pragma once

#ifdef __cplusplus
extern "C"
#endif
void treecode_force_mrag(const float theta,
			 const float * const xsources,
			 const float * const ysources,
			 const float * const sourcevalues,
			 const int nsources,
			 const float * const x0s,
			 const float * const y0s,
			 const float * const hs,
			 const int nblocks,
			 float * const xresults,
			 float * const yresults);

/*
 * 2d-treecodes-force.h -- drop-in header for treecode_force_mrag.
 * Same declaration as the reference's generated header
 * (libraries.Makefile:35-36 from drivers/treecode-force.h:21-31, realtype=double).
 * Implemented by libtreecodes2d.so (also installed as lib2d-treecodes-force.so).
 * Output layout: the result for point (ix, iy) of block ib is at ix + 32 * (iy + 32 * ib).
 */
#pragma once

#ifdef __cplusplus
extern "C"
#endif
void treecode_force_mrag(const double theta,
			 const double * const xsources,
			 const double * const ysources,
			 const double * const sourcevalues,
			 const int nsources,
			 const double * const x0s,
			 const double * const y0s,
			 const double * const hs,
			 const int nblocks,
			 double * const xresults,
			 double * const yresults);

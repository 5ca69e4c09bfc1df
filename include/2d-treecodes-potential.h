/*
 * 2d-treecodes-potential.h -- drop-in header for treecode_potential.
 * Same declaration as the reference's generated header
 * (libraries.Makefile:30-31 from drivers/treecode-potential.h:21-29, realtype=double).
 * Implemented by libtreecodes2d.so (also installed as lib2d-treecodes-potential.so).
 */
#pragma once

#ifdef __cplusplus
extern "C"
#endif
void treecode_potential(const double theta,
			const double * const xsources,
			const double * const ysources,
			const double * const sourcevalues,
			const int nsources,
			const double * const xtargets,
			const double * const ytargets,
			const int ntargets,
			double * const targetvalues);

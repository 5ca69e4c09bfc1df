/*
 * 2d-treecodes-potential.h (float) -- drop-in header of the reference's real=float build.
 * Same declaration as the reference's generated header
 * (libraries.Makefile:30-31 from drivers/treecode-potential.h:21-29, realtype=float).
 * Implemented by libtreecodes2d_float.so (also installed as
 * 2d-treecodes_amd/float/lib2d-treecodes-potential.so): FP64 tree geometry and
 * translations, FP32 near field and multipole evaluation.
 */
#pragma once

#ifdef __cplusplus
extern "C"
#endif
void treecode_potential(const float theta,
			const float * const xsources,
			const float * const ysources,
			const float * const sourcevalues,
			const int nsources,
			const float * const xtargets,
			const float * const ytargets,
			const int ntargets,
			float * const targetvalues);

// gpucell/compat.hpp -- header-only drop-in for GPUCell-style host code.
//
// Declares namespaces cupar and cudiv with the reference's exact parameter
// lists (src/ParticleGPU.cuh:8-37, src/ConfigDomain.cuh:9-18), taking the
// HIP vector types the reference's host code passes (float3*, float4*,
// int2*, int3*).  Each function forwards to the C ABI of libgpucell_hip.so
// and, like the reference's CHECK macro (src/erro.cuh:6-20), prints the
// failure and exits the process when a call fails.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "gpucell.h"

static_assert(sizeof(float3) == sizeof(tfloat3), "float3 / tfloat3");
static_assert(sizeof(float4) == sizeof(tfloat4), "float4 / tfloat4");
static_assert(sizeof(int2) == sizeof(tint2), "int2 / tint2");
static_assert(sizeof(int3) == sizeof(tint3), "int3 / tint3");

namespace gpucell_compat {
inline void check(int status, const char* fn, const char* file, int line)
{
    if (status != 0) {
        std::printf("GPUCell error:\n    File:       %s\n    Line:       %d\n"
                    "    Call:       %s\n    Status:     %d\n    Error text: %s\n",
                    file, line, fn, status, gc_GetLastError());
        std::exit(1);
    }
}
template <class T, class U> inline T* as(U* p) { return reinterpret_cast<T*>(p); }
}  // namespace gpucell_compat

#define GPUCELL_CALL(fn, ...) gpucell_compat::check(fn(__VA_ARGS__), #fn, __FILE__, __LINE__)

namespace cudiv {
using gpucell_compat::as;
inline void PreConList(int n, int cutradius, float3* d_totalcoor, float* d_bboxsph, float* d_invwid,
                       int* d_ncube, int3* d_cc, int* d_cubepart, int* d_sortpart, int* d_oldpos)
{
    GPUCELL_CALL(gc_PreConList, n, cutradius, as<tfloat3>(d_totalcoor), d_bboxsph, d_invwid, d_ncube,
                 as<tint3>(d_cc), d_cubepart, d_sortpart, d_oldpos);
}
inline void Sort(int n, int* keys, int* values) { GPUCELL_CALL(gc_Sort, n, keys, values); }
inline void CalBeginEndCube(int n, int SizeCube, int* d_cubepart, int2* d_beginendcube)
{
    GPUCELL_CALL(gc_CalBeginEndCube, n, SizeCube, d_cubepart, as<tint2>(d_beginendcube));
}
inline void SortBasicArrays(int n, int* d_sortpart, int* d_oldpos, float3* d_totalcoor, int3* d_cc,
                            float4* d_velrhop, float3* d_velhalf, float3* d_totalforce, float* d_press)
{
    GPUCELL_CALL(gc_SortBasicArrays, n, d_sortpart, d_oldpos, as<tfloat3>(d_totalcoor), as<tint3>(d_cc),
                 as<tfloat4>(d_velrhop), as<tfloat3>(d_velhalf), as<tfloat3>(d_totalforce), d_press);
}
inline void SortArraysWithCells(int n, int* d_sortpart, int* d_oldpos, int* d_curpos, float3* d_totalcoor,
                                int3* d_cc, float4* d_velrhop, float3* d_velhalf, float3* d_totalforce,
                                float* d_press, int3* d_nthrobound)
{
    GPUCELL_CALL(gc_SortArraysWithCells, n, d_sortpart, d_oldpos, d_curpos, as<tfloat3>(d_totalcoor),
                 as<tint3>(d_cc), as<tfloat4>(d_velrhop), as<tfloat3>(d_velhalf),
                 as<tfloat3>(d_totalforce), d_press, as<tint3>(d_nthrobound));
}
}  // namespace cudiv

namespace cupar {
using gpucell_compat::as;
inline void ConstsDataUp(const Structconst* consts) { GPUCELL_CALL(gc_ConstsDataUp, consts); }
inline void UpdateFluidCoor(int n, int* d_sortpart, float3* d_Totalforce, float3* d_totalcoor,
                            float4* d_velrhop)
{
    GPUCELL_CALL(gc_UpdateFluidCoor, n, d_sortpart, as<tfloat3>(d_Totalforce), as<tfloat3>(d_totalcoor),
                 as<tfloat4>(d_velrhop));
}
inline void AppFluidPeriodicBC(int n, int* d_sortpart, float3* d_totalcoor, int3* d_nthrobound,
                               float* d_bboxsdpd)
{
    GPUCELL_CALL(gc_AppFluidPeriodicBC, n, d_sortpart, as<tfloat3>(d_totalcoor), as<tint3>(d_nthrobound),
                 d_bboxsdpd);
}
inline void CalFluidDen(int n, float3* d_totalcoor, float4* d_velrhop, int3* d_cc, int* d_ncube,
                        float* d_outerregion, int* d_sortpart, int2* d_beginendcube)
{
    GPUCELL_CALL(gc_CalFluidDen, n, as<tfloat3>(d_totalcoor), as<tfloat4>(d_velrhop), as<tint3>(d_cc),
                 d_ncube, d_outerregion, d_sortpart, as<tint2>(d_beginendcube));
}
inline void PredictFluidVel(int n, int* d_sortpart, float4* d_velrhop, float3* d_velhalf,
                            float3* d_Totalforce)
{
    GPUCELL_CALL(gc_PredictFluidVel, n, d_sortpart, as<tfloat4>(d_velrhop), as<tfloat3>(d_velhalf),
                 as<tfloat3>(d_Totalforce));
}
inline void CalBound2Vel(int n, float3* d_totalcoor, float4* d_velrhop, int3* d_cc, int* d_ncube,
                         float* d_outerregion, int* d_sortpart, int2* d_beginendcube)
{
    GPUCELL_CALL(gc_CalBound2Vel, n, as<tfloat3>(d_totalcoor), as<tfloat4>(d_velrhop), as<tint3>(d_cc),
                 d_ncube, d_outerregion, d_sortpart, as<tint2>(d_beginendcube));
}
inline void EOS(int n, int* d_sortpart, float4* d_velrhop, float* d_press)
{
    GPUCELL_CALL(gc_EOS, n, d_sortpart, as<tfloat4>(d_velrhop), d_press);
}
inline void CalFluidSDPDForce(int n, tfloat3* d_sdpdforce, float3* d_totalcoor, float4* d_velrhop,
                              int3* d_cc, int* d_ncube, int* d_sortpart, float* d_outerreigon,
                              int2* d_beginendcube, float3* d_totalforce, float* d_press)
{
    GPUCELL_CALL(gc_CalFluidSDPDForce, n, d_sdpdforce, as<tfloat3>(d_totalcoor), as<tfloat4>(d_velrhop),
                 as<tint3>(d_cc), d_ncube, d_sortpart, d_outerreigon, as<tint2>(d_beginendcube),
                 as<tfloat3>(d_totalforce), d_press);
}
inline void CalFluidRepForce(int n, float3* d_totalcoor, int3* d_cc, int* d_ncube, int* d_sortpart,
                             float* d_outerreigon, int2* d_beginendcube, float3* d_totalforce)
{
    GPUCELL_CALL(gc_CalFluidRepForce, n, as<tfloat3>(d_totalcoor), as<tint3>(d_cc), d_ncube, d_sortpart,
                 d_outerreigon, as<tint2>(d_beginendcube), as<tfloat3>(d_totalforce));
}
inline void SetFluidExtForce(int DriveOn, int n, int* d_sortpart, float3* d_totalforce)
{
    GPUCELL_CALL(gc_SetFluidExtForce, DriveOn, n, d_sortpart, as<tfloat3>(d_totalforce));
}
inline void CorrectFliudVel(int n, int* d_sortpart, float3* d_totalforce, float3* d_velhalf,
                            float4* d_velrhop, float3* d_totalcoor, int3* d_cc, int* d_ncube,
                            float* d_outerreigon, int2* d_beginendcube)
{
    GPUCELL_CALL(gc_CorrectFliudVel, n, d_sortpart, as<tfloat3>(d_totalforce), as<tfloat3>(d_velhalf),
                 as<tfloat4>(d_velrhop), as<tfloat3>(d_totalcoor), as<tint3>(d_cc), d_ncube,
                 d_outerreigon, as<tint2>(d_beginendcube));
}
inline void CalGeoPara(long stepcount, int ntri, int ncell, float* d_bbox, int* d_curpos, int3* d_tri,
                       int2* d_beginendtri, int3* d_nthrobound, float3* d_totalcoor,
                       float3* d_trinordir, float* d_triarea, float* d_reftriarea, float3* d_tricenter,
                       float* d_refarea_para, float* d_refvol_para, float* d_area_para,
                       float* d_vol_para, float* d_area, float* d_vol, float* d_pararea)
{
    GPUCELL_CALL(gc_CalGeoPara, stepcount, ntri, ncell, d_bbox, d_curpos, as<tint3>(d_tri),
                 as<tint2>(d_beginendtri), as<tint3>(d_nthrobound), as<tfloat3>(d_totalcoor),
                 as<tfloat3>(d_trinordir), d_triarea, d_reftriarea, as<tfloat3>(d_tricenter),
                 d_refarea_para, d_refvol_para, d_area_para, d_vol_para, d_area, d_vol, d_pararea);
}
inline void SetMechPara(int ncell, int nedge, int6* d_edge, int* d_curpos, float* d_refedgelen,
                        float* d_maxrefedgelen, float* d_bendmodpara, float* d_shearmodpara,
                        float* d_hbending, float* d_powerindex, float* d_perlenWLC, float* d_hpow,
                        float* d_reftriangle, float3* d_trinordir, float3* d_tricente, float3* d_coor)
{
    GPUCELL_CALL(gc_SetMechPara, ncell, nedge, d_edge, d_curpos, d_refedgelen, d_maxrefedgelen,
                 d_bendmodpara, d_shearmodpara, d_hbending, d_powerindex, d_perlenWLC, d_hpow,
                 d_reftriangle, as<tfloat3>(d_trinordir), as<tfloat3>(d_tricente), as<tfloat3>(d_coor));
}
inline void CalCellDefForce(int ncell, unsigned nedge, unsigned ntri, unsigned ntotalpar,
                            unsigned ncellpar, float* d_bbox, int6* d_edge, int3* d_tri, float3* d_coor,
                            tfloat3* defforce, float3* d_defforce, int3* d_nthrobound,
                            float* d_maxrefedgelen, float* d_powerindex, float* d_perlenWLC,
                            float* d_hpow, float3* d_trinordir, float3* d_tricenter,
                            float* d_reftriangle, float* d_hbending, float* d_triarea,
                            float* d_reftriarea, float* d_hlarea_para, float* d_hgarea_para,
                            float* d_area_para, float* d_refarea_para, float* d_hvol_para,
                            float* d_vol_para, float* d_refvol_para, float3* d_totalforce,
                            int* d_sortpart, int* d_curpos)
{
    GPUCELL_CALL(gc_CalCellDefForce, ncell, nedge, ntri, ntotalpar, ncellpar, d_bbox, d_edge,
                 as<tint3>(d_tri), as<tfloat3>(d_coor), defforce, as<tfloat3>(d_defforce),
                 as<tint3>(d_nthrobound), d_maxrefedgelen, d_powerindex, d_perlenWLC, d_hpow,
                 as<tfloat3>(d_trinordir), as<tfloat3>(d_tricenter), d_reftriangle, d_hbending,
                 d_triarea, d_reftriarea, d_hlarea_para, d_hgarea_para, d_area_para, d_refarea_para,
                 d_hvol_para, d_vol_para, d_refvol_para, as<tfloat3>(d_totalforce), d_sortpart,
                 d_curpos);
}
inline void CalCellAggForce(unsigned ncell, unsigned n, unsigned ncellpar, float3* d_totalcoor,
                            float3* d_aggforce, int3* d_cc, int* d_ncube, int* d_sortpart, int* d_celltab,
                            float* d_outerreigon, int2* d_beginendcube, float3* d_totalforce,
                            float* d_pararea)
{
    GPUCELL_CALL(gc_CalCellAggForce, ncell, n, ncellpar, as<tfloat3>(d_totalcoor),
                 as<tfloat3>(d_aggforce), as<tint3>(d_cc), d_ncube, d_sortpart, d_celltab,
                 d_outerreigon, as<tint2>(d_beginendcube), as<tfloat3>(d_totalforce), d_pararea);
}
}  // namespace cupar

#undef GPUCELL_CALL

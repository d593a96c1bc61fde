// Colour data of the spectral path: the CIE 1931 2-degree standard observer
// and the CIE D65 relative spectral power distribution, 360 .. 830 nm at
// 1 nm (471 rows).  Public standard data (CIE 015:2018; CIE S 014-1 and
// S 014-2).  The reference renders with the same numbers
// (Core/ColorFunctions.h:64-87, Core/ColorFunctions.cpp), so this table
// holds exactly its values: image parity needs the same observer and
// illuminant on both sides.  Plain C++ (no HIP), so the oracle can compile it.
#pragma once

namespace mray
{

constexpr int kCieRows = 471;            // 360 .. 830 nm
constexpr int kCieFirstNm = 360;

// Row i: xbar, ybar, zbar and the D65 SPD (relative, 100 at 560 nm) at
// 360 + i nm.
extern const float kCieObserverD65[kCieRows][4];

// Trapezoid integrals of the observer over the table (1 nm steps), the
// factors each channel is normalised by before it is uploaded
// (Tracer/SpectrumContext.cu:401-426 with Core/ColorFunctions.h:70-72).
constexpr float kCieIntegralX = 106.86540391770f;
constexpr float kCieIntegralY = 106.85691492064f;
constexpr float kCieIntegralZ = 106.89194822846f;
// Integral of D65 times ybar (the factor of rgb2spec the reference uses,
// Core/ColorFunctions.h:85): the illuminant is scaled by
// kCieIntegralY / kD65NormFactor before upload.
constexpr float kD65NormFactor = 10566.864f;
// Rec.709 primaries and the D65 white point the reference uses for
// MR_REC_709 (Core/ColorFunctions.h:57, :405-414).
constexpr double kRec709Primaries[8] = {0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.31272, 0.32903};

} // namespace mray

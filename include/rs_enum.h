/*
 * rs_enum.h -- slot meanings of the float16 / float4 descriptors and the
 * simulation concept bits.  The numeric values are the reference's
 * (rs_enum.h:1-137): descriptors are exchanged with callers and stored in the
 * state dump, so every slot index must mean what it means there.
 */
#ifndef RS_ENUM_H
#define RS_ENUM_H

enum RS_GPU_VENDOR {
    RS_GPU_VENDOR_UNKNOWN,
    RS_GPU_VENDOR_NVIDIA,
    RS_GPU_VENDOR_INTEL,
    RS_GPU_VENDOR_AMD
};

/* 1-D table: index = clamp(x * scale + origin, 0, maximum) */
enum RSTable1DDescription {
    RSTable1DDescriptionScale        = 0,
    RSTable1DDescriptionOrigin       = 1,
    RSTable1DDescriptionMaximum      = 2,
    RSTable1DDescriptionUserConstant = 3
};

/* Uniform 3-D table: coordinate = fma(position, scale, origin) */
enum RSTable3DDescription {
    RSTable3DDescriptionScaleX = 0, RSTable3DDescriptionScaleY = 1, RSTable3DDescriptionScaleZ = 2,
    RSTable3DDescriptionRefreshTime = 3,
    RSTable3DDescriptionOriginX = 4, RSTable3DDescriptionOriginY = 5, RSTable3DDescriptionOriginZ = 6,
    RSTable3DDescriptionFormat = 7,
    RSTable3DDescriptionMaximumX = 8, RSTable3DDescriptionMaximumY = 9, RSTable3DDescriptionMaximumZ = 10,
    RSTable3DDescription11 = 11,
    RSTable3DDescriptionRecipInLnX = 12, RSTable3DDescriptionRecipInLnY = 13, RSTable3DDescriptionRecipInLnZ = 14,
    RSTable3DDescriptionTachikawa = 15
};

/* Stretched 3-D table: coordinate = copysign(base, p) log1p(scale |p|) + offset */
enum RSTable3DStaggeredDescription {
    RSTable3DStaggeredDescriptionBaseChangeX = 0, RSTable3DStaggeredDescriptionBaseChangeY = 1,
    RSTable3DStaggeredDescriptionBaseChangeZ = 2,
    RSTable3DStaggeredDescriptionRefreshTime = 3,
    RSTable3DStaggeredDescriptionPositionScaleX = 4, RSTable3DStaggeredDescriptionPositionScaleY = 5,
    RSTable3DStaggeredDescriptionPositionScaleZ = 6,
    RSTable3DStaggeredDescriptionFormat = 7,
    RSTable3DStaggeredDescriptionOffsetX = 8, RSTable3DStaggeredDescriptionOffsetY = 9,
    RSTable3DStaggeredDescriptionOffsetZ = 10,
    RSTable3DStaggeredDescription11 = 11,
    RSTable3DStaggeredDescriptionRecipInLnX = 12, RSTable3DStaggeredDescriptionRecipInLnY = 13,
    RSTable3DStaggeredDescriptionRecipInLnZ = 14,
    RSTable3DStaggeredDescriptionTachikawa = 15
};

enum RSTableDescription {
    RSTableDescriptionScaleX = 0, RSTableDescriptionScaleY = 1, RSTableDescriptionScaleZ = 2,
    RSTableDescriptionReserved1 = 3,
    RSTableDescriptionOriginX = 4, RSTableDescriptionOriginY = 5, RSTableDescriptionOriginZ = 6,
    RSTableDescriptionReserved2 = 7,
    RSTableDescriptionMaximumX = 8, RSTableDescriptionMaximumY = 9, RSTableDescriptionMaximumZ = 10,
    RSTableDescriptionReserved3 = 11, RSTableDescriptionReserved4 = 12, RSTableDescriptionReserved5 = 13,
    RSTableDescriptionReserved6 = 14, RSTableDescriptionReserved7 = 15
};

enum RSSimulationDescription {
    RSSimulationDescriptionBeamUnitX = 0,
    RSSimulationDescriptionBeamUnitY = 1,
    RSSimulationDescriptionBeamUnitZ = 2,
    RSSimulationDescriptionTotalParticles = 3,
    RSSimulationDescriptionWaveNumber = 4,
    RSSimulationDescriptionConcept = 5,
    RSSimulationDescriptionDropConcentrationScale = 6,
    RSSimulationDescriptionSimTic = 7,
    RSSimulationDescriptionBoundOriginX = 8,
    RSSimulationDescriptionBoundOriginY = 9,
    RSSimulationDescriptionBoundOriginZ = 10,
    RSSimulationDescriptionPRT = 11,
    RSSimulationDescriptionBoundSizeX = 12,
    RSSimulationDescriptionBoundSizeY = 13,
    RSSimulationDescriptionBoundSizeZ = 14,
    RSSimulationDescription15 = 15
};

enum RSDropSizeDistribution {
    RSDropSizeDistributionUndefined = 0,
    RSDropSizeDistributionMarshallPalmer = 1,
    RSDropSizeDistributionGamma = 2,
    RSDropSizeDistributionArbitrary = 3
};

enum RSTableSpacing {
    RSTableSpacingUniform = 0,
    RSTableSpacingStretchedX = 1,
    RSTableSpacingStretchedY = 1 << 1,
    RSTableSpacingStretchedZ = 1 << 2,
    RSTableSpacingStretchedXYZ = 7
};

enum RSSimulationConcept {
    RSSimulationConceptNull = 0,
    RSSimulationConceptNonZeroCrossPol = 1,
    RSSimulationConceptDraggedBackground = 1 << 1,
    RSSimulationConceptTransparentBackground = 1 << 2,
    RSSimulationConceptBoundedParticleVelocity = 1 << 3,
    RSSimulationConceptUniformDSDScaledRCS = 1 << 4,
    RSSimulationConceptFixedScattererPosition = 1 << 5,
    RSSimulationConceptVerticallyPointingRadar = 1 << 6,
    RSSimulationConceptDebrisFluxFromVelocity = 1 << 7
};

#endif

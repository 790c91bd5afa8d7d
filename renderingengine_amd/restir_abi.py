"""numpy views of the ReSTIR DI records the library keeps in HBM: the
reference's versioned ABI structs (include/contracts/RestirAbiV3.hpp:109-193,
include/contracts/ReconstructionAbiV2.hpp's GpuPrimarySurfaceV2)."""
from __future__ import annotations

import numpy as np

_F4 = (np.float32, 4)
_U4 = (np.uint32, 4)

LIGHT_SAMPLE = np.dtype([("positionDistance", *_F4), ("directionCombinedPdf", *_F4),
                         ("radianceDiscretePdf", *_F4), ("conditionalPdf", *_F4), ("identity", *_U4),
                         ("generation", *_U4), ("metadata", *_U4), ("sourceIdentity", *_U4)])
RESERVOIR = np.dtype([("selected", LIGHT_SAMPLE), ("weightState", *_F4), ("selectedTerms", *_F4),
                      ("state", *_U4), ("provenance", *_U4), ("surfaceIdentity", *_U4), ("sceneIdentity", *_U4),
                      ("frameIdentity", *_U4), ("reprojection", *_U4)])
SURFACE = np.dtype([("worldPositionLinearDepth", *_F4), ("geometricNormalRoughness", *_F4),
                    ("shadingNormalMetallic", *_F4), ("diffuseAlbedo", *_F4), ("specularAlbedo", *_F4),
                    ("identity", *_U4)])
STATISTICS = np.dtype([("candidateCounts", *_U4), ("reuseCounts", *_U4), ("visibilityCounts", *_U4),
                       ("errorMetrics", *_F4)])
assert (LIGHT_SAMPLE.itemsize, RESERVOIR.itemsize, SURFACE.itemsize, STATISTICS.itemsize) == (128, 256, 96, 64)

# Reservoir flags (state.z) and rejection reasons (state.w), RestirAbiV3.hpp:58-101.
VALID, M_CLAMPED, TEMPORAL_ACCEPTED, SPATIAL_ACCEPTED = 1, 2, 4, 8
VISIBILITY_EVALUATED, VISIBILITY_VALID = 32, 64
REJECTIONS = ("none", "invalid-candidate", "reprojection-outside", "motion-invalid", "camera-cut", "resize", "depth",
              "normal", "instance", "thin-geometry", "scene-generation", "light-generation", "age", "material")

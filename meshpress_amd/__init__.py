"""MI355X-native STRIDE (DPR4 v3) mesh codec hot path.

Surfaces (drop-ins for maletsden/meshpress):
  meshpress_amd.encoder.STRIDEEncoder
  meshpress_amd.utils.paradelta_v5_cuda.ParaDeltaV5GpuDecoder
  meshpress_amd.utils.types.Model / CompressedModel
  meshpress_amd.batch.BatchDecoder   (many streams: one launch of each decode pass,
                                     or one hipGraph replay of them)
"""
__version__ = "0.1.0"

"""SIFT front end and prior-guided descriptor matcher of FieldStitch / PlaneSite for MI355X
(gfx950): hand-written HIP kernels in libfsx.so behind a C ABI (include/fsx.h), called through
ctypes; torch tensors serve as device buffers only.

Modules (each a drop-in for the reference module named):

* ``gsift``      detect / describe (app/planesite/gsift.py)
* ``features``   sift_gpu, _spread, _save, the batched FeatureBatch and the GPU worker loop
                 (app/planesite/features.py)
* ``matching``   _get, _ratio_match_bf, _unique, _guided, _predict, match_pair and the batched
                 PairMatcher / FloatMatcher (app/planesite/matching.py)
* ``camera``     undistortion (app/planesite/camera.py); ``pairs`` the pipeline's GPS pair
                 selection (app/planesite/pipeline.py)
* ``verify``     deterministic GPU RANSAC (homography, fundamental matrix, similarity)
* ``clahe``      createCLAHE(...).apply, the photos' input conditioning (app/planesite/features.py)
* ``tracks``     build_tracks: feature tracks from verified match lists (app/planesite/recon.py)
* ``filter_mlp`` the learned BF16 match filter; ``autotune`` launch-shape tuning; ``synth``
                 seeded synthetic flights; ``build`` compiles libfsx.so in-tree
"""

"""gpupolyaurnlda_amd -- MI355X-native Polya-urn LDA Gibbs sampler.

A drop-in HIP replacement (gfx950) for the CUDA library behind the reference's five-call C ABI
(src/ffi.rs:15-31): ``initialize``, ``sample_phi``, ``sample_z_async``, ``sync_buffer``,
``cleanup``, plus extension entry points (seed, state read-back, device-resident corpus, stage
kernels).  Python modules:

* ``train``     -- ``Sampler`` over the five calls and the train.rs buffer loop;
* ``resident``  -- ``ResidentCorpus``, the whole corpus in HBM (throughput form: schedules,
                   hipGraph replay, autotuner, log-likelihood);
* ``serve``     -- ``FoldInServer``, continuous-batching fold-in against the resident model;
* ``stages``    -- per-kernel entry points on numpy arrays;
* ``corpus``    -- seeded synthetic corpora and serving requests;
* ``calibrate`` -- a plain copy kernel for scaling hardware byte counters;
* ``build``     -- the hipcc / g++ build recipe (library, host helpers, ``gpulda_train`` driver).
"""
__all__ = ["train", "resident", "serve", "stages", "corpus", "calibrate", "build"]

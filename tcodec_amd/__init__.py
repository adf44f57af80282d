"""tcodec_amd - MI355X-native motion estimation for the TCodec video codec.

* ``tcodec_amd.me``     ctypes binding of libtcme.so (drop-in ``tc_motion_est``
  and the batched ``tcme_*`` API)
* ``tcodec_amd.synth``  seeded synthetic video generators for the benchmark
  workloads
* ``tcodec_amd.workload`` call sets for throughput runs (quadtree nodes x
  references x centres)
"""
__all__ = ["me", "synth", "workload"]

"""simbody_amd: MI355X-native batched tree dynamics (Simbody's tree path).

The compute path is the HIP library libtdyn.so (simbody_amd/csrc); this
package holds its host-side Python mirror (simbody_amd.matter) and the
workload generators (simbody_amd.models).
"""

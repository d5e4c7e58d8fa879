"""MI355X engine for PaStiX numerical factorisation and solve.

`pastix_amd.pastix` binds the PaStiX C API of libpastix_amd.so;
`pastix_amd.spm` builds spmatrix_t descriptors and the synthetic workloads.
"""

"""Test oracle for the SPECT projector path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product (niftyrec_amd, lib_et_array_interface.so) never
does.
"""

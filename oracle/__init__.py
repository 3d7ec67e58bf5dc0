"""CPU oracle of the LES_Halo timestep - TEST INFRASTRUCTURE ONLY.

A sequential C++ restatement of the reference Fortran physics (oracle/les_oracle.cpp), used as
the checker by tests/, by the smoke check and by the benchmark's CPU-baseline leg.  The product
library (les_halo_amd) never imports, links or calls anything under oracle/.
"""

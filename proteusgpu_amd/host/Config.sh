# Compile-time options; the Makefile turns every uncommented line into a -D flag
# (the reference's Config.sh convention).

# 2D or 3D (one must be defined)
#dim_2D
dim_3D

# HDF5 for IC and snapshot files
USE_HDF5

# Dump each mesh build's KNN problem to output_knn_file (HDF5)
#WRITE_KNN_OUTPUT

# Recompute seeds that fail the security-radius certificate at a larger k
# (comment out for the reference's behaviour: such cells stay failed)
CERT_RESTART

# HLL face orientation: outward normals (comment out for the reference's)
#PHYSICAL_FLUX

# Dry run: print the banner and exit
#DRY_RUN

# Neighbours per seed and per-cell capacities
_K_=32
_MAX_P_=64
_MAX_T_=96

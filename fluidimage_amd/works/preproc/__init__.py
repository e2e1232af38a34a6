"""Preprocessing of PIV images on the GPU (the reference's
``fluidimage.works.preproc`` sliding filters)."""

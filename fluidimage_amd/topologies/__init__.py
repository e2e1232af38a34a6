"""Topologies: series of image files through the GPU works."""

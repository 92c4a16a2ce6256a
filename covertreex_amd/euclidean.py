"""Euclidean cover tree on the GPU (reference Rust Euclidean path,
src/algo.rs:179-323 and src/lib.rs:207-223).  Not built yet: raises."""

from __future__ import annotations


def euclidean_cover_tree(coords, *, chunk_size: int = 512):
    raise NotImplementedError("the Euclidean HIP cover tree is not available in this build")

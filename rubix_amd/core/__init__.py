"""Drop-in step API: the rubix.core step factories of the particle ->
datacube path, each `get_<step>(config) -> Callable[[RubixData], RubixData]`
with the reference's config keys, error messages and RubixData fields,
running hand-written HIP kernels on torch device tensors."""

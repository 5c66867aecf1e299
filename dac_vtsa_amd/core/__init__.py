"""Drop-in core: TVSAModel (residual), FDJacobian, TRBDF2 over the HIP C ABI."""

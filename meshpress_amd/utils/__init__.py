"""Host-side mirrors of the reference's `utils` surfaces on the STRIDE path."""

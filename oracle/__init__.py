"""CPU FP64 oracle: test infrastructure and CPU baseline only (never imported by
the product package)."""

"""ORACLE package — test infrastructure only (CPU restatement used as the
checker and the CPU baseline; never imported by the product)."""

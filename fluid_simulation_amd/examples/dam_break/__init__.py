"""Dam-break scene and driver (counterparts of the reference's examples/dam_break)."""

"""FARM-MI355X: the polygon-join hot path of tengdj/FARM (raster filter,
segment refinement, intersection area, within distance) on gfx950.

farm_amd.capi binds libfarm.so (include/farm_capi.h, include/farm_host.h);
farm_amd.datagen builds the seeded synthetic polygon sets.
"""
__all__ = ["capi", "datagen"]

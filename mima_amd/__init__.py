"""MI355X (gfx950) clear-sky RRTMG column radiation for MiMA.

``mima_amd.rrtmg`` is the Python face of the C ABI in include/rrtmg_hip.h;
``mima_amd.columns`` generates seeded synthetic GCM columns.
"""

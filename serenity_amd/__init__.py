"""MI355X-native XC Fock build (the FuncPotential path of Serenity) for gfx950.

``serenity_amd.xc`` holds the Serenity-shaped host classes over the C ABI of
``libserenity_xc.so`` (include/serenity_xc.h); ``serenity_amd.workloads``
builds the seeded synthetic systems of the benchmark configurations.
"""
__version__ = "0.1.0"

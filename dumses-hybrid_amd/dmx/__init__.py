"""Host side of the MI355X-native DUMSES-Hybrid MHD step.

dmx.abi       ctypes binding of libdmx.so (the C ABI of include/dmx.h)
dmx.problems  initial states on the reference grid
dmx.dumpfile  the reference's binary dump/restart format
dmx.namelist  the reference's namelist input file
"""

"""Host side of the MI355X-native DUMSES-Hybrid MHD step.

dmx.abi             ctypes binding of libdmx.so (the C ABI of include/dmx.h)
dmx.problems        initial states on the reference grid
dmx.dumpfile        the reference's binary dump/restart format
dmx.driver          the reference time loop's history cadence over the ABI
dmx.learned         float64 reference of the learned slope limiter
dmx.train_limiter   training and evaluation of the learned limiter
dmx.shear_precision float32 against float64 on the shearing-box MRI
dmx.rocprof_summary summaries of rocprofv3 output
"""

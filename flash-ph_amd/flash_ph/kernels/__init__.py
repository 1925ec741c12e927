"""Stand-alone kernel entry points kept at the reference's module paths
(flash_ph.kernels.*) for callers that drive single stages."""

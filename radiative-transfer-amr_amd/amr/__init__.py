"""Error indicators of the adaptive loop (``src/amr``): the analytic L2
error (``Error_Indicator.error_analytic``) on the GPU."""

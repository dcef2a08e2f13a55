"""Test and baseline infrastructure: the FP64 oracle (tomocam_oracle) and the
FP32 CPU restatement used as the benchmark's CPU baseline (cpu_port).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package; the product (tomocam_amd) never imports it."""

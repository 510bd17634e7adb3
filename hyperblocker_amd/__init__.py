"""hyperblocker_amd: MI355X (gfx950) rule-based blocking for entity resolution.

A drop-in for HyperBlocker's blocking hot path: the same YAML rules, CSV
inputs, execution-plan encoding and left_id,right_id output, evaluated by
HIP kernels through the C ABI in include/hyperblocker_amd.h.
"""
from .blocker import (ArrayPlan, ColumnTable, Context, DeviceBytes, DevicePlan,
                      DeviceSerializedTable, DeviceTable, HostPlan, HostTable, Job, SerializedTable,
                      Server, filter_coefs, partition, sort_pairs, write_pairs)
from .capi import HBError

__all__ = ["ArrayPlan", "ColumnTable", "Context", "DeviceBytes", "DevicePlan",
           "DeviceSerializedTable", "DeviceTable", "HostPlan", "Server",
           "HostTable", "Job", "HBError", "SerializedTable", "filter_coefs", "partition", "sort_pairs",
           "write_pairs"]

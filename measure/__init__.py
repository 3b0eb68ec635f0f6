"""Measurement code beside bench.py: the benchmark modes other than the default assembly line
(batched, serving, uv, Newton) and the rocprofv3 summary helper. Not part of the product
package; the CPU-baseline legs here are the only place that runs the oracle."""

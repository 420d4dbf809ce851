"""CPU oracle for the path-tracing hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package, and only as the checker (never as the measured or shipped path).
"""

"""CPU oracle of the SnuQS state-vector path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, as the checker; the product (snuqs_amd) never imports it.
"""

"""The learned per-tile cost model of the balanced pipeline form.

model.py  the cost_model.bin format (KzCostModel) and numpy references of
          its inference arithmetic
train.py  training data from profiled plain-form frames and training with
          the driver's MFMA training kernel (kazan_cost_model_train)
"""

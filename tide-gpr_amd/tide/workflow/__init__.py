"""Workflow helpers around the operators: shot batching and receiver
objectives (reference ``src/tide/workflow``)."""

from .losses import (
    LossNormalization,
    backward_shot_batches,
    receiver_mse_loss,
    take_receiver_batch,
)
from .objective import ReceiverLoss, ReceiverObjective
from .shots import (
    ShotBatch,
    batch_operator,
    infer_receiver_shot_dim,
    index_shots,
    merge_receiver_batches,
    run_shot_batches,
    split_shots,
    take_shot_batch,
)

__all__ = [
    "LossNormalization",
    "ReceiverLoss",
    "ReceiverObjective",
    "ShotBatch",
    "backward_shot_batches",
    "batch_operator",
    "infer_receiver_shot_dim",
    "index_shots",
    "merge_receiver_batches",
    "receiver_mse_loss",
    "run_shot_batches",
    "split_shots",
    "take_receiver_batch",
    "take_shot_batch",
]

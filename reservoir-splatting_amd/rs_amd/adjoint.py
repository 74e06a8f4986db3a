"""torch.autograd.Function over the frozen-reservoir replay (rs_replay_forward /
rs_replay_backward): the image as a differentiable function of the materials' base colour
and roughness, with the pass's last frame frozen (ReservoirSplatting.freeze())."""
import torch


class FrozenReplay(torch.autograd.Function):
    """R = FrozenReplay.apply(params, rs, height, width): params is a (numMaterials, 4)
    float32 CUDA tensor (base colour RGB, roughness); R is (height, width, 4) float32."""

    @staticmethod
    def forward(ctx, params, rs, height, width):
        params = params.contiguous()
        image = torch.empty((height, width, 4), dtype=torch.float32, device=params.device)
        rs.replay_forward(params, image, torch.cuda.current_stream())
        ctx.rs = rs
        ctx.save_for_backward(params)
        return image

    @staticmethod
    def backward(ctx, grad_image):
        (params,) = ctx.saved_tensors
        grad = torch.empty(params.shape, dtype=torch.float64, device=params.device)
        ctx.rs.replay_backward(params, grad_image.contiguous().float(), grad, torch.cuda.current_stream())
        return grad.to(params.dtype), None, None, None


def replay(rs, params, height, width):
    return FrozenReplay.apply(params, rs, height, width)
